// pb_noalloc.hip -- the device scratch of the allocation-free velocity entry
// point (pb_velocity_noalloc): one file-scope device array of PB_NOALLOC_BYTES,
// held by every device the library runs on until the process ends.
//
// In host code the array's name is only a host-side stand-in: its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current, at pb_create) and kept here; host code uses only that address.
#include "pb_internal.h"

#include <mutex>

namespace {

static __device__ __align__(256) unsigned char scratch[PB_NOALLOC_BYTES];

constexpr int kMaxDevices = 64;
void *scratch_ptr[kMaxDevices] = {};
bool scratch_done[kMaxDevices] = {};
std::mutex scratch_mutex;

} // namespace

int pb_noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return PB_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_done[device]) return scratch_ptr[device] ? PB_OK : PB_ERR_HIP;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return PB_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return PB_ERR_HIP;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    scratch_done[device] = true;
    scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    return scratch_ptr[device] ? PB_OK : PB_ERR_HIP;
}

void *pb_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    return scratch_ptr[device];
}
