// pxr_runtime.h -- host-side runtime shared by the C-ABI entry points:
// current stream, error reporting, the device error counter and a small
// cache of library-owned scratch buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pxrhip.h"

namespace pxr {

struct Runtime {
  int device = -1;
  hipStream_t stream = nullptr;
  bool own_stream = false;  // created by pxrhip_init
  bool capturing = false;   // a graph capture of the stream is open
  bool checks = true;
  unsigned long long *d_bad = nullptr;  // device counter of skipped particles
  int num_cus = 256;
  // leaf deposition summed in LDS blocks (pxr_leafbin.hip) from this many
  // particles on (0: never), work items of at most leaf_bin_chunk particles
  int64_t leaf_bin_min = 1 << 16;
  int leaf_bin_chunk = 4096;
  // incremented whenever library-owned device memory that recorded work may
  // point at is freed (a scratch slot that grows, the cached tile table):
  // a HIP graph captured before that refers to freed memory
  unsigned long long mem_gen = 0;
};

Runtime &rt();
int set_error(int code, const char *fmt, ...);
int ensure_init();
// library-owned scratch buffer number `slot`, grown to at least `bytes`
void *scratch(int slot, size_t bytes);
// read the bad-particle counter (synchronises the stream), reset it to 0
int take_bad(unsigned long long *n);

inline unsigned grid1d(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace pxr

#define PXR_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return pxr::set_error(PXRHIP_ERR_HIP, "%s failed: %s", #call,           \
                            hipGetErrorString(e_));                            \
  } while (0)

#define PXR_LAUNCHED()                                                         \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess)                                                      \
      return pxr::set_error(PXRHIP_ERR_HIP, "kernel launch failed: %s",        \
                            hipGetErrorString(e_));                            \
  } while (0)

#define PXR_INIT()                                                             \
  do {                                                                         \
    int r_ = pxr::ensure_init();                                               \
    if (r_ != PXRHIP_OK) return r_;                                            \
  } while (0)
