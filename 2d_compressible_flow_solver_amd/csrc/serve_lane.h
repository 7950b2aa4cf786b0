// serve_lane.h -- library-internal facts of a batch that the serving loop
// (serve.cpp) needs to route a request past the batch (batch.hip).
#pragma once
#include <stdint.h>

#include "../../include/cfs.h"

// Hidden: visible to the library's own files, not exported from libcfs.so.
#define CFS_INTERNAL __attribute__((visibility("hidden")))

extern "C" {
// CFS_OK when cfs_batch_admit would accept the problem given room: the
// checks that do not depend on free arena space or slots.
CFS_INTERNAL int32_t batch_check_problem(const cfs_batch* b, const cfs_config* cfg, const cfs_immersed* ib,
                                         double cfl, double t_end);
// Tiles of the problem in the batch's tile shape, and the number of
// workgroups the GPU holds at once for the batch's stage kernel.
CFS_INTERNAL int64_t batch_problem_tiles(const cfs_batch* b, const cfs_config* cfg);
CFS_INTERNAL int64_t batch_resident_workgroups(const cfs_batch* b);
// Device memory free on the batch's device now (hipMemGetInfo).
CFS_INTERNAL int32_t batch_device_free_bytes(const cfs_batch* b, int64_t* bytes);
// Makes the batch's device current on the calling thread (a worker thread
// starts on the default device).
CFS_INTERNAL int32_t batch_make_device_current(const cfs_batch* b);
}
