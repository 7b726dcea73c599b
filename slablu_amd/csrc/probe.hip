// Launch probe: HIP events around every launch of one chosen kernel, on the
// stream that kernel is launched on, with each launch's algorithmic flops
// and bytes (fixed formulas of its shape).  It times one kernel inside a
// larger run without a profiler: the sum of the event intervals over the
// probed launches, divided by their count, is the kernel's average launch
// duration.  Launches into a stream under graph capture are not probed
// (their replays are invisible to the host).  One probe at a time per
// process; not for concurrent callers.
#include "slab_internal.h"
#include "../../include/slab_hip.h"

#include <atomic>
#include <mutex>
#include <vector>

namespace slab {
namespace {

struct Probe {
  std::mutex mu;
  int kernel = 0;
  int cap = 0, used = 0, dropped = 0;
  std::vector<hipEvent_t> ev;  // start/end event pairs, created at start
  double flops = 0, bytes = 0;
};
Probe g_probe;
std::atomic<int> g_probe_kernel{0};

}  // namespace

bool probe_on(int kernel) { return g_probe_kernel.load(std::memory_order_relaxed) == kernel; }

int probe_begin(int kernel, hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return -1;
  std::lock_guard<std::mutex> lock(g_probe.mu);
  if (g_probe.kernel != kernel) return -1;
  if (g_probe.used >= g_probe.cap) {
    ++g_probe.dropped;
    return -1;
  }
  const int i = g_probe.used++;
  if (hipEventRecord(g_probe.ev[2 * i], s) != hipSuccess) {
    --g_probe.used;
    ++g_probe.dropped;
    return -1;
  }
  return i;
}

void probe_end(int slot, hipStream_t s, double flops, double bytes) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lock(g_probe.mu);
  hipEventRecord(g_probe.ev[2 * slot + 1], s);
  g_probe.flops += flops;
  g_probe.bytes += bytes;
}

}  // namespace slab

using namespace slab;

extern "C" int slab_probe_start(int kernel, int max_launches) {
  if (kernel < SLAB_PROBE_GEMM_F32 || kernel > SLAB_PROBE_GEMM_F64 || max_launches < 1 || max_launches > (1 << 20))
    return -1;
  std::lock_guard<std::mutex> lock(g_probe.mu);
  if (g_probe.kernel != 0) return -3;  // a probe is running
  for (hipEvent_t e : g_probe.ev) hipEventDestroy(e);
  g_probe.ev.assign(2 * (size_t)max_launches, nullptr);
  for (auto& e : g_probe.ev) {
    const hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      for (hipEvent_t x : g_probe.ev)
        if (x) hipEventDestroy(x);
      g_probe.ev.clear();
      return -(1000 + (int)r);
    }
  }
  g_probe.cap = max_launches;
  g_probe.used = g_probe.dropped = 0;
  g_probe.flops = g_probe.bytes = 0;
  g_probe.kernel = kernel;
  g_probe_kernel.store(kernel);
  return 0;
}

extern "C" int slab_probe_stop(int* launches, int* dropped, double* total_ms, double* flops, double* bytes) {
  std::lock_guard<std::mutex> lock(g_probe.mu);
  if (g_probe.kernel == 0) return -1;
  g_probe_kernel.store(0);
  g_probe.kernel = 0;
  double ms = 0;
  for (int i = 0; i < g_probe.used; ++i) {
    hipError_t r = hipEventSynchronize(g_probe.ev[2 * i + 1]);
    float t = 0;
    if (r == hipSuccess) r = hipEventElapsedTime(&t, g_probe.ev[2 * i], g_probe.ev[2 * i + 1]);
    if (r != hipSuccess) return -(1000 + (int)r);
    ms += t;
  }
  if (launches) *launches = g_probe.used;
  if (dropped) *dropped = g_probe.dropped;
  if (total_ms) *total_ms = ms;
  if (flops) *flops = g_probe.flops;
  if (bytes) *bytes = g_probe.bytes;
  return 0;
}
