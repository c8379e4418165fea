// pxr_runtime.hip -- library set-up, stream and error handling.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "pxr_runtime.h"

namespace pxr {

static Runtime g_rt;
static thread_local char g_err[1024] = "";

struct Scratch {
  void *p = nullptr;
  size_t bytes = 0;
};
static std::vector<Scratch> g_scratch(32);

Runtime &rt() { return g_rt; }

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int ensure_init() {
  if (g_rt.device >= 0) return PXRHIP_OK;
  return pxrhip_init(-1);
}

void *scratch(int slot, size_t bytes) {
  if (slot < 0 || slot >= (int)g_scratch.size()) return nullptr;
  Scratch &s = g_scratch[slot];
  if (s.bytes >= bytes && s.p) return s.p;
  if (s.p) {
    hipStreamSynchronize(g_rt.stream);
    hipFree(s.p);
    s.p = nullptr;
    s.bytes = 0;
    g_rt.mem_gen++;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (hipMalloc(&s.p, want) != hipSuccess) {
    s.p = nullptr;
    return nullptr;
  }
  s.bytes = want;
  return s.p;
}

int take_bad(unsigned long long *n) {
  unsigned long long h = 0;
  PXR_HIP(hipMemcpyAsync(&h, g_rt.d_bad, sizeof(h), hipMemcpyDeviceToHost, g_rt.stream));
  PXR_HIP(hipStreamSynchronize(g_rt.stream));
  if (h) PXR_HIP(hipMemsetAsync(g_rt.d_bad, 0, sizeof(h), g_rt.stream));
  *n = h;
  return PXRHIP_OK;
}

}  // namespace pxr

using namespace pxr;

extern "C" {

int pxrhip_init(int device) {
  if (g_rt.device >= 0 && (device < 0 || device == g_rt.device)) return PXRHIP_OK;
  int dev = device;
  if (dev < 0) PXR_HIP(hipGetDevice(&dev));
  PXR_HIP(hipSetDevice(dev));
  hipDeviceProp_t prop;
  PXR_HIP(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(PXRHIP_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950",
                     dev, prop.gcnArchName);
  g_rt.num_cus = prop.multiProcessorCount;
  if (!g_rt.stream) {
    // the library's own stream (default priority, ordered with the null
    // stream); it can be captured into a HIP graph, the null stream cannot
    PXR_HIP(hipStreamCreate(&g_rt.stream));
    g_rt.own_stream = true;
  }
  if (!g_rt.d_bad) {
    PXR_HIP(hipMalloc(&g_rt.d_bad, 2 * sizeof(unsigned long long)));
    PXR_HIP(hipMemset(g_rt.d_bad, 0, 2 * sizeof(unsigned long long)));
  }
  g_rt.device = dev;
  return PXRHIP_OK;
}

int pxrhip_finalize(void) {
  if (g_rt.device < 0) return PXRHIP_OK;
  hipDeviceSynchronize();
  for (auto &s : g_scratch) {
    if (s.p) hipFree(s.p);
    s.p = nullptr;
    s.bytes = 0;
  }
  if (g_rt.d_bad) hipFree(g_rt.d_bad);
  g_rt.d_bad = nullptr;
  if (g_rt.own_stream && g_rt.stream) hipStreamDestroy(g_rt.stream);
  g_rt.stream = nullptr;
  g_rt.own_stream = false;
  g_rt.device = -1;
  return PXRHIP_OK;
}

const char *pxrhip_last_error(void) { return g_err; }

int pxrhip_set_stream(void *s) {
  if (g_rt.capturing) return set_error(PXRHIP_ERR_ARG, "the stream is being captured");
  g_rt.stream = (hipStream_t)s;
  g_rt.own_stream = false;
  return PXRHIP_OK;
}

// ---- HIP graphs of the library's work ---------------------------------------
struct GraphHandle {
  hipGraph_t graph;
  hipGraphExec_t exec;
  unsigned long long mem_gen;  // Runtime::mem_gen when it was captured
};
static unsigned long long g_capture_gen = 0;

int pxrhip_graph_begin(void) {
  PXR_INIT();
  if (!g_rt.stream) return set_error(PXRHIP_ERR_ARG, "the null stream cannot be captured");
  if (g_rt.capturing) return set_error(PXRHIP_ERR_ARG, "a capture is already open");
  PXR_HIP(hipStreamBeginCapture(g_rt.stream, hipStreamCaptureModeRelaxed));
  g_rt.capturing = true;
  g_capture_gen = g_rt.mem_gen;
  return PXRHIP_OK;
}

int pxrhip_graph_end(void **handle) {
  PXR_INIT();
  if (!handle) return set_error(PXRHIP_ERR_ARG, "null handle");
  *handle = nullptr;
  if (!g_rt.capturing) return set_error(PXRHIP_ERR_ARG, "no capture is open");
  g_rt.capturing = false;
  hipGraph_t g = nullptr;
  PXR_HIP(hipStreamEndCapture(g_rt.stream, &g));
  if (g_rt.mem_gen != g_capture_gen) {
    hipGraphDestroy(g);
    return set_error(PXRHIP_ERR_ARG, "library memory was reallocated while the graph was "
                     "recorded; run the work once before capturing it");
  }
  hipGraphExec_t e = nullptr;
  hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  if (err != hipSuccess) {
    hipGraphDestroy(g);
    return set_error(PXRHIP_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(err));
  }
  *handle = new GraphHandle{g, e, g_capture_gen};
  return PXRHIP_OK;
}

int pxrhip_graph_launch(void *handle) {
  PXR_INIT();
  if (!handle) return set_error(PXRHIP_ERR_ARG, "null graph");
  GraphHandle *h = (GraphHandle *)handle;
  if (h->mem_gen != g_rt.mem_gen)
    return set_error(PXRHIP_ERR_ARG, "the graph refers to library memory that has been "
                     "reallocated since it was captured (a scratch buffer grew or the tile "
                     "table changed); capture it again");
  PXR_HIP(hipGraphLaunch(h->exec, g_rt.stream));
  return PXRHIP_OK;
}

int pxrhip_graph_destroy(void *handle) {
  if (!handle) return PXRHIP_OK;
  GraphHandle *h = (GraphHandle *)handle;
  if (g_rt.stream) (void)hipStreamSynchronize(g_rt.stream);
  (void)hipGraphExecDestroy(h->exec);
  (void)hipGraphDestroy(h->graph);
  delete h;
  return PXRHIP_OK;
}

void *pxrhip_get_stream(void) { return (void *)g_rt.stream; }

int pxrhip_synchronize(void) {
  PXR_HIP(hipStreamSynchronize(g_rt.stream));
  return PXRHIP_OK;
}

int pxrhip_set_checks(int on) {
  g_rt.checks = on != 0;
  return PXRHIP_OK;
}

int pxrhip_version(void) { return 100; }

int pxrhip_bad_count(long long *count, int reset) {
  PXR_INIT();
  unsigned long long h = 0;
  PXR_HIP(hipMemcpyAsync(&h, g_rt.d_bad, sizeof(h), hipMemcpyDeviceToHost, g_rt.stream));
  PXR_HIP(hipStreamSynchronize(g_rt.stream));
  if (reset) PXR_HIP(hipMemsetAsync(g_rt.d_bad, 0, sizeof(h), g_rt.stream));
  if (count) *count = (long long)h;
  return PXRHIP_OK;
}

}  // extern "C"
