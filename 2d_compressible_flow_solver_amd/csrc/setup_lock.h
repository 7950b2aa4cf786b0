// setup_lock.h -- HIP error reporting and the serialisation of the library's
// stream captures with its allocation and set-up calls, shared by every
// source of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "../../include/cfs.h"

namespace cfs {

// A failed HIP call: CFS_E_HIP, and with CFS_DEBUG set in the environment the
// call, its place and the HIP error on stderr.
inline int hip_failed(hipError_t e, const char* call, const char* file, int line) {
    const char* dbg = std::getenv("CFS_DEBUG");
    if (dbg && *dbg) std::fprintf(stderr, "cfs: %s failed at %s:%d: %s\n", call, file, line, hipGetErrorName(e));
    return CFS_E_HIP;
}

// The library's stream captures and its allocations, frees, event and stream
// set-up are made one at a time across threads (the serving loop's workers
// create, free and capture beside the batch's own capture): the runtime
// synchronises streams inside some of these calls, which must not meet a
// capture in progress on another thread.  Recursive: cfs_create cleans up
// through cfs_destroy.
inline std::recursive_mutex& hip_setup_mutex() {
    static std::recursive_mutex m;
    return m;
}
using SetupLock = std::lock_guard<std::recursive_mutex>;

}  // namespace cfs

#define HIP_OK(x)                                                                               \
    do {                                                                                        \
        const hipError_t hip_ok_e_ = (x);                                                       \
        if (hip_ok_e_ != hipSuccess) return cfs::hip_failed(hip_ok_e_, #x, __FILE__, __LINE__); \
    } while (0)
