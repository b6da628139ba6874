// ge3d.cpp -- host side of the 3D gradient-echo train: the block timing, the line table, the
// timing checks, and the batched partition transform's descriptors (include/mrsim_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "internal.h"
#include "tables.h"

namespace mrs {
hipError_t launch_zstage(int mt, bool adjoint, const ZSet* sets, const ZBlock* blocks, int nblocks, hipStream_t st);
hipError_t launch_partition_noise(float2* k, int nz, int C, int ny, int nx, const int32_t* rows, int nrows, uint64_t g,
                                  uint64_t seed, int n_spikes, double amp, double sigma, hipStream_t st);
int zstage_columns_per_block();
}  // namespace mrs

using namespace mrs;

namespace {

int fail(char* err, size_t errlen, int code, const char* fmt, ...) {
    if (err && errlen) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

}  // namespace

// ---------------------------------------------------------------------------- partition transform

namespace mrs {

// Validates the public sets and lays out the device descriptors and block list, grouped by table
// size: blocks of table class k (nz <= 16, 32, 64) are [off[k], off[k+1]).
int build_zsets(const mrs_zset* in, size_t n, bool adjoint, std::vector<ZSet>& sets, std::vector<ZBlock>& blocks,
                int off[4], char* err, size_t errlen) {
    sets.clear();
    blocks.clear();
    if (n && !in) return fail(err, errlen, MRS_E_INVALID, "null z-stage sets");
    std::vector<ZBlock> per[3];
    const int span = zstage_columns_per_block();
    for (size_t i = 0; i < n; i++) {
        const mrs_zset& z = in[i];
        if (!z.in || !z.out) return fail(err, errlen, MRS_E_INVALID, "z-stage set %zu: null buffer", i);
        if ((const void*)z.in == (const void*)z.out)
            return fail(err, errlen, MRS_E_INVALID, "z-stage set %zu: the transform is not in place", i);
        if (z.nz < 1 || z.nz > 64)
            return fail(err, errlen, MRS_E_UNSUPPORTED, "z-stage set %zu: %u partitions (1..64 supported)", i, z.nz);
        if (z.n_comp < 1 || z.n_coils < 1 || z.ny < 1 || z.nx < 1 || z.ky_segments < 1)
            return fail(err, errlen, MRS_E_INVALID, "z-stage set %zu: empty dimension", i);
        if (z.ny % z.ky_segments != 0)
            return fail(err, errlen, MRS_E_INVALID, "%u phase-encode lines do not divide into %u ky segments", z.ny,
                        z.ky_segments);
        const size_t ncols = (size_t)z.n_coils * z.ny * z.nx;
        if (ncols * z.nz >= (size_t)1 << 31)
            return fail(err, errlen, MRS_E_UNSUPPORTED, "z-stage set %zu: more than 2^31 samples", i);
        if (z.n_comp > 1 && z.comp_stride < ncols * z.nz)
            return fail(err, errlen, MRS_E_INVALID, "z-stage set %zu: compartment blocks overlap (stride %zu < %zu)",
                        i, z.comp_stride, ncols * z.nz);
        ZSet s{};
        s.in = reinterpret_cast<const float2*>(z.in);
        s.out = reinterpret_cast<float2*>(z.out);
        s.w = z.weights;
        s.cstride = (int64_t)z.comp_stride;
        s.nz = (int32_t)z.nz;
        s.ncols = (int32_t)ncols;
        s.ny = (int32_t)z.ny;
        s.nx = (int32_t)z.nx;
        s.kseg = (int32_t)z.ky_segments;
        s.nq = (int32_t)z.n_comp;
        s.accumulate = z.accumulate ? 1 : 0;
        s.mt = z.nz <= 16 ? 1 : (z.nz <= 32 ? 2 : 4);
        const int cls = s.mt == 1 ? 0 : (s.mt == 2 ? 1 : 2);
        for (size_t c = 0; c < ncols; c += (size_t)span) per[cls].push_back(ZBlock{(int32_t)i, (int32_t)c});
        sets.push_back(s);
    }
    (void)adjoint;
    off[0] = 0;
    for (int k = 0; k < 3; k++) {
        blocks.insert(blocks.end(), per[k].begin(), per[k].end());
        off[k + 1] = (int)blocks.size();
    }
    return MRS_OK;
}

hipError_t enqueue_zsets(const ZSet* d_sets, const ZBlock* d_blocks, const int off[4], bool adjoint, hipStream_t st) {
    static const int mts[3] = {1, 2, 4};
    for (int k = 0; k < 3; k++) {
        const int nb = off[k + 1] - off[k];
        if (!nb) continue;
        hipError_t e = launch_zstage(mts[k], adjoint, d_sets, d_blocks + off[k], nb, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace mrs

extern "C" int mrs_zstage(const mrs_zset* in, size_t nsets, int32_t adjoint, void* stream, char* err, size_t errlen) {
    std::vector<ZSet> sets;
    std::vector<ZBlock> blocks;
    int off[4];
    int rc = build_zsets(in, nsets, adjoint != 0, sets, blocks, off, err, errlen);
    if (rc || sets.empty()) return rc;
    hipStream_t st = (hipStream_t)stream;
    void* mem = nullptr;
    const size_t sb = sets.size() * sizeof(ZSet), bb = blocks.size() * sizeof(ZBlock);
    hipError_t e = hipMalloc(&mem, sb + bb);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
    ZSet* d_sets = (ZSet*)mem;
    ZBlock* d_blocks = (ZBlock*)((char*)mem + sb);
    if ((e = hipMemcpyAsync(d_sets, sets.data(), sb, hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = hipMemcpyAsync(d_blocks, blocks.data(), bb, hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = enqueue_zsets(d_sets, d_blocks, off, adjoint != 0, st)) == hipSuccess)
        e = hipStreamSynchronize(st);
    hipFree(mem);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "z-stage: %s", hipGetErrorString(e));
    return MRS_OK;
}

// A z-stage whose descriptors live on the device: created once (one blocking upload), then run
// asynchronously as often as needed, with no allocation or wait per run.
struct mrs_zplan {
    std::vector<mrs::ZSet> sets;
    int off[4];
    bool adjoint;
    void* mem = nullptr;
    mrs::ZSet* d_sets = nullptr;
    mrs::ZBlock* d_blocks = nullptr;
};

extern "C" int mrs_zplan_create(const mrs_zset* in, size_t nsets, int32_t adjoint, mrs_zplan** out, char* err,
                                size_t errlen) {
    if (!out) return fail(err, errlen, MRS_E_INVALID, "null output handle");
    *out = nullptr;
    std::vector<mrs::ZSet> sets;
    std::vector<mrs::ZBlock> blocks;
    int off[4];
    int rc = mrs::build_zsets(in, nsets, adjoint != 0, sets, blocks, off, err, errlen);
    if (rc) return rc;
    mrs_zplan* z = new mrs_zplan;
    z->adjoint = adjoint != 0;
    std::copy(off, off + 4, z->off);
    const size_t sb = sets.size() * sizeof(mrs::ZSet), bb = blocks.size() * sizeof(mrs::ZBlock);
    if (sb + bb > 0) {
        hipError_t e = hipMalloc(&z->mem, sb + bb);
        if (e != hipSuccess) {
            delete z;
            return fail(err, errlen, MRS_E_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
        }
        z->d_sets = (mrs::ZSet*)z->mem;
        z->d_blocks = (mrs::ZBlock*)((char*)z->mem + sb);
        if ((e = hipMemcpy(z->d_sets, sets.data(), sb, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(z->d_blocks, blocks.data(), bb, hipMemcpyHostToDevice)) != hipSuccess) {
            hipFree(z->mem);
            delete z;
            return fail(err, errlen, MRS_E_DEVICE, "z-plan upload: %s", hipGetErrorString(e));
        }
    }
    z->sets.swap(sets);
    *out = z;
    return MRS_OK;
}

extern "C" int mrs_zplan_run(mrs_zplan* z, void* stream, char* err, size_t errlen) {
    if (!z) return fail(err, errlen, MRS_E_INVALID, "null z-plan");
    if (z->sets.empty()) return MRS_OK;
    const hipError_t e = mrs::enqueue_zsets(z->d_sets, z->d_blocks, z->off, z->adjoint, (hipStream_t)stream);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "z-stage: %s", hipGetErrorString(e));
    return MRS_OK;
}

extern "C" void mrs_zplan_destroy(mrs_zplan* z) {
    if (!z) return;
    if (z->mem) hipFree(z->mem);
    delete z;
}

// partitions_to_slices' spikes and noise (src/kspace3d.rs:235-264) on one volume's acquired 3D k-space
// [p][coil][ky][kx] (device), for volume index `volume` on `seed` (the echo's salted seed).
extern "C" int mrs_partition_noise(mrs_c32* kspace, size_t nz, size_t n_coils, size_t nx, size_t ny,
                                   const mrs_acquisition* acq, uint64_t volume, uint64_t seed, void* stream, char* err,
                                   size_t errlen) {
    if (!kspace || !acq) return fail(err, errlen, MRS_E_INVALID, "null argument");
    const double sigma = acq->noise_variance > 0.0 ? std::sqrt(acq->noise_variance / (double)(nx * ny)) : 0.0;
    if (acq->n_spikes == 0 && sigma == 0.0) return MRS_OK;
    std::vector<uint8_t> mask(ny);
    mrs_sampling_mask(1, ny, acq, mask.data());
    std::vector<int32_t> rows;
    for (size_t ky = 0; ky < ny; ky++)
        if (mask[ky]) rows.push_back((int32_t)ky);
    int32_t* drows = nullptr;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMalloc(&drows, std::max<size_t>(rows.size(), 1) * 4);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(drows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_partition_noise(reinterpret_cast<float2*>(kspace), (int)nz, (int)n_coils, (int)ny, (int)nx, drows,
                                   (int)rows.size(), volume, seed, (int)acq->n_spikes, acq->spike_amplitude, sigma, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(drows);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "partition noise: %s", hipGetErrorString(e));
    return MRS_OK;
}
