/*
 * launch_serve.h -- host-side launchers of the serving-pool kernels.
 */
#pragma once

#include <hip/hip_runtime.h>

#include "serve.h"

namespace gray {

hipError_t launch_srv_init(const SrvRequest *reqs, const SrvSlots &S, const int *a_slot,
                           const int *a_req, const int *a_ray, int n, hipStream_t st);
/* which: 0 FP32 Kerr, 1 FP64 Kerr, 2 FP32 grid */
hipError_t launch_srv_epoch(int which, const SrvRequest *reqs, const SrvSlots &S, const int *list,
                            int n, int K, const SrvLists &out, hipStream_t st);

} // namespace gray
