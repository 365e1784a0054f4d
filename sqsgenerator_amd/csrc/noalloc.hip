// sqsg_run_noalloc: a run that allocates no device memory.  Its threshold, counters and
// record sink live in `scratch`, a file-scope device array of fixed size that every GPU the
// library runs on holds until the process ends.  In host code the array's name is only a
// stand-in for a device address, so the address of each device's copy is looked up once
// per device with hipGetSymbolAddress and used everywhere instead of the name.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/sqsgen_amd.h"
#include "job.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mutex;                      // one call at a time per process (they share it)
void *g_scratch[kMaxDevices] = {nullptr};  // each device's copy of `scratch`

void *device_scratch() {
  int dev = 0;
  SQSG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::out_of_range("device index out of range");
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    SQSG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw sqsg::DeviceError("no device address for the scratch buffer");
    g_scratch[dev] = ptr;
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" int sqsg_run_noalloc(sqsg_job *job, sqsg_progress_fn cb, void *user) {
  return sqsg::guarded([&] {
    if (!job) throw std::invalid_argument("job must not be NULL");
    std::lock_guard<std::mutex> lock(g_mutex);
    sqsg::FixedScratch fs;
    fs.base = device_scratch();
    fs.bytes = kScratchBytes;
    sqsg::run_job(job, cb, user, &fs);
  });
}
