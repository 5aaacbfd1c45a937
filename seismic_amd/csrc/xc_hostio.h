// Host <-> device copies of caller-owned (pageable) buffers through pinned
// staging, for the calls that take host buffers (the drop-ins and
// sxc_allpairs_host*).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace sxc {
// One array of a multi-array copy: `bytes` at `host` (pageable) and `dev`.
struct HostSeg {
  void* host;
  void* dev;
  size_t bytes;
};
// Copies several arrays (host -> device, device -> host) as one pipeline of
// staged chunks, in the order given, on `st`; returns when all are in place.
// SXC_OK, or an error code with the error string set.
int copy_segs_to_device(const HostSeg* segs, int n_segs, hipStream_t st);
int copy_segs_to_host(const HostSeg* segs, int n_segs, hipStream_t st);
// Copies `bytes` from pageable host memory to device memory on `st` and
// returns when the copy is complete.  The data goes through two pinned
// staging buffers: the host-side copy of one chunk (several threads) runs
// while the DMA of the previous one is in flight.  SXC_OK, or an error code
// with the error string set.
int copy_to_device(void* dst, const void* src, size_t bytes, hipStream_t st);
// The same for device memory -> pageable host memory: the DMA of the next
// chunk runs while the previous one is copied out of its staging buffer.
int copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t st);
}  // namespace sxc
