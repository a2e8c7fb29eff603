// Shared helpers of the device library's translation units: the thread-local error text behind
// tt_last_error, status checks, and the stream cast.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <initializer_list>
#include <string>

#include "../../../include/tentorch_amd.h"

namespace ttdev {

// Records `message` as this thread's last error and returns the failure status.
int fail(const std::string &message);

// Counts one leg-move or block-product kernel launch (read back by tt_launch_count).
void note_launch();

// Kernel timing (tt_timing_enable): when on for this thread, returns a fresh start/stop event pair
// for the next launch and remembers it; otherwise both are null.
struct TimedPair {
  hipEvent_t start = nullptr, stop = nullptr;
};
int timed_pair(TimedPair *pair);

// Launches `kernel` on `stream`, with the kernel-dispatch events of timing mode when it is on.
template <typename Kernel, typename... Args>
int launch(Kernel kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t stream, Args... args) {
  TimedPair pair;
  if (int rc = timed_pair(&pair)) return rc;
  if (pair.start)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, pair.start, pair.stop, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, stream, args...);
  return 0;
}

// Several host tables of one plan copied to the device together (each at a 256-byte aligned
// offset): *block receives their device range (handed back with release_tables), each
// *table.where its table's address (null when empty).  The range is carved from a device arena
// of large chunks (no allocation per plan) and filled by an asynchronous copy from a pinned ring
// on the library's own upload stream, which the host does not wait for: every launch that reads
// plan tables calls await_uploads on its stream first.
struct Table {
  void **where;
  const void *data;
  size_t bytes;
};
int upload_tables(void **block, std::initializer_list<Table> tables);

// Returns a plan's table range to the arena; a chunk that no live plan uses any more, and that
// takes no new tables, is freed.
void release_tables(void *block);

// Orders `stream` after every table upload issued so far on the current device: nothing when
// they have completed, else a device-side wait on the upload stream's event (a host wait for the
// uploads when `stream` is being captured into a graph).
int await_uploads(hipStream_t stream);

inline hipStream_t as_stream(void *stream) { return reinterpret_cast<hipStream_t>(stream); }

struct cdouble {
  double re, im;
};

}  // namespace ttdev

#define TT_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t status_ = (call);                                                               \
    if (status_ != hipSuccess)                                                                 \
      return ::ttdev::fail(std::string(#call) + " failed: " + hipGetErrorString(status_));     \
  } while (0)

#define TT_LAUNCHED()                                                                          \
  do {                                                                                         \
    hipError_t status_ = hipGetLastError();                                                    \
    if (status_ != hipSuccess)                                                                 \
      return ::ttdev::fail(std::string("kernel launch failed: ") + hipGetErrorString(status_)); \
  } while (0)
