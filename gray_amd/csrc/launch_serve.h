/*
 * launch_serve.h -- host-side launchers of the serving pool's epoch kernels
 * (kernels_serve.hip).
 */
#pragma once

#include <hip/hip_runtime.h>

#include "serve.h"

namespace gray {

hipError_t launch_srv_clear(SrvCtr *ctr, unsigned long long budget, hipStream_t st);
hipError_t launch_srv_kerr(const SrvKerrArgs &A, int n_blocks, hipStream_t st);
hipError_t launch_srv_fb64(const SrvFbArgs &A, int n_blocks, hipStream_t st);
hipError_t launch_srv_grid(const SrvGridArgs &A, int n_blocks, hipStream_t st);
/* Resident blocks per CU of the Kerr (256 threads) and grid (4 waves)
 * kernels. */
hipError_t srv_kerr_blocks_per_cu(int *blocks);
hipError_t srv_grid_blocks_per_cu(int *blocks);
/* Kerr blocks a CU holds beside one FP64 or grid block. */
hipError_t srv_kerr_beside_heavy(int *blocks);

} // namespace gray
