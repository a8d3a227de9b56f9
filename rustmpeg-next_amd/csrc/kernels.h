// kernels.h - launch interface between the C-ABI host layer (rmh.cpp) and the
// HIP kernels (residual.hip, mc.hip, deblock.hip). Internal to librmh.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rmh.h"

namespace rmh {

// Device-resident constant tables, built once per context.
struct DeviceTables {
    const uint32_t* mat = nullptr;   // transform matrices as int16 pairs: [kk][j] = (M[2kk][j], M[2kk+1][j])
    const int8_t* lfnst8 = nullptr;  // [4][2][16][48]
    const int8_t* lfnst4 = nullptr;  // [4][2][16][16]
    const uint16_t* scans = nullptr; // grouped diagonal scans, raster positions
    int32_t mat_off[3][7];           // dword offset of (type, log2 n); -1 if unsupported
    int32_t scan_off[7][7];          // element offset of scan (log2 w, log2 h)
    int32_t scan_len[7][7];
};

struct ResidualArgs {
    const rmh_tu_desc* tus;
    const int16_t* coeff;
    int16_t* residual;
    const uint32_t* mat;
    const int8_t* lfnst8;
    const int8_t* lfnst4;
    const uint16_t* scans;
    uint32_t first_tu;
    uint32_t n_tu;
    int32_t mat_off[3][7];
    int32_t scan_off[7][7];
    int32_t scan_len[7][7];
};

hipError_t launch_residual(const ResidualArgs& a, hipStream_t stream);

// One inter work item: a 16x16 luma tile (and its 8x8 chroma tiles) of one PU.
struct McItem {
    uint32_t pu;
    uint16_t tx;  // tile column inside the PU (units of 16 luma samples)
    uint16_t ty;  // tile row
};

hipError_t launch_mc(const rmh_mc_frame* frames, const rmh_pu_desc* pus, const McItem* items, uint32_t n_items,
                     hipStream_t stream);

// Deblocking runs two launches (all vertical edges, then all horizontal edges).
hipError_t launch_deblock(const rmh_dbk_frame* frames, uint32_t n_frames, uint32_t max_cells, hipStream_t stream);

}  // namespace rmh
