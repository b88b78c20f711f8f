// The allocation-free decode entry point: prores_decode_device_batch's
// device-planned batch with its whole workspace in `scratch`, a file-scope
// device array of NOALLOC_SCRATCH_BYTES that each GPU holds from the first
// use of this library on it until the process ends. Set-up
// (prores_noalloc_init) loads the library's kernels and constant tables and
// its stream on the current device and looks up that device's copy of
// `scratch`; after it, a batch allocates no device or page-locked memory. The
// host staging of jobs and results is a static array per device as well.
#include <hip/hip_runtime.h>

#include "host_internal.h"

using namespace prores_hip;

namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = 256ull << 20;
constexpr size_t NOALLOC_HOST_BYTES = 256u << 10;  // jobs up and results back: room for NOALLOC_MAX_JOBS
constexpr size_t NOALLOC_MAX_JOBS = 2048;
constexpr int MAX_DEVICES = 64;
}  // namespace

// One copy per GPU, at its own address; host code names it only through the
// address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
struct NoallocDevice {
  void *ptr = nullptr;  // this device's `scratch`
  Device *dev = nullptr;
  Lane *lane = nullptr;
};
std::mutex noalloc_mu;
NoallocDevice noalloc_devices[MAX_DEVICES];
alignas(256) uint8_t noalloc_host[MAX_DEVICES][NOALLOC_HOST_BYTES];

Err current(int &id) {
  const hipError_t e = hipGetDevice(&id);
  if (e != hipSuccess) return device_error("hipGetDevice", e);
  if (id < 0 || id >= MAX_DEVICES) return Err{PRORES_ERR_DEVICE, "device index out of range"};
  return Err{};
}
}  // namespace

extern "C" {

size_t prores_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

prores_status prores_noalloc_init(prores_error *err) {
  int id = 0;
  Err e = current(id);
  Device *dev = nullptr;
  Lane *L = nullptr;
  if (e.ok()) e = device(dev);
  if (e.ok()) e = lane(*dev, NOALLOC_LANE, L);
  if (e.ok()) {
    std::lock_guard<std::mutex> g(noalloc_mu);
    NoallocDevice &nd = noalloc_devices[id];
    if (!nd.ptr) {
      void *ptr = nullptr;
      e = check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
      if (e.ok() && !ptr) e = Err{PRORES_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave no address"};
      if (e.ok()) {
        std::lock_guard<std::mutex> lg(L->mu);
        (void)decode_workspace(*L);  // the lane's workspace record, made now
        nd.ptr = ptr;
        nd.dev = dev;
        nd.lane = L;
      }
    }
  }
  fill_error(err, e);
  return (prores_status)e.kind;
}

prores_status prores_decode_device_batch_noalloc(const prores_decoder *d, size_t n, const prores_dev_decode_job *jobs,
                                                 prores_error *errs) {
  int id = 0;
  Err e = current(id);
  NoallocDevice nd;
  if (e.ok()) {
    std::lock_guard<std::mutex> g(noalloc_mu);
    nd = noalloc_devices[id];
    if (!nd.ptr) e = config_error("prores_noalloc_init has not run on this device");
  }
  if (e.ok() && n > NOALLOC_MAX_JOBS) e = config_error("a batch of more than 2048 jobs for the scratch memory");
  if (!e.ok()) {
    for (size_t i = 0; i < n; i++) fill_error(&errs[i], e);
    return (prores_status)e.kind;
  }
  // Calls on one device share its scratch: they run one at a time.
  std::lock_guard<std::mutex> g(nd.lane->mu);
  FixedRegion region;
  region.dev = (uint8_t *)nd.ptr;
  region.dev_bytes = NOALLOC_SCRATCH_BYTES;
  region.host = noalloc_host[id];
  region.host_bytes = NOALLOC_HOST_BYTES;
  e = decode_batch_fixed(*nd.dev, *nd.lane, d, n, jobs, region, errs);
  return e.ok() ? PRORES_OK : (prores_status)e.kind;
}

}  // extern "C"
