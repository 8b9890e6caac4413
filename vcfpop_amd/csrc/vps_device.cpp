// vps_device.cpp — plain device memory and device control behind the C ABI,
// for callers that keep buffers of their own on the device (the reference's
// MallocDeviceCUDA / FreeCUDA / MemsetCUDA / MemcpyCUDA / WaitCUDA /
// ResetDeviceCUDA / ShowDevicesCUDA, src/structureCUDA.h:8-22, 38-42).
// Unlike the reference's checkErrors (src/structureCUDA.cpp:183-196), every
// failure is a status code.
#include <string.h>
#include "vps_internal.h"

using namespace vps;

extern "C" int vps_device_alloc(int device, size_t bytes, void** out)
{
    if (!out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    VPS_HIP(hipSetDevice(device));
    const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(VPS_ERR_NOMEM, "device %d: %zu bytes: %s", device, bytes, hipGetErrorString(e));
    }
    return VPS_OK;
}

extern "C" int vps_device_free(void* p)
{
    if (p) VPS_HIP(hipFree(p));
    return VPS_OK;
}

extern "C" int vps_device_memset(void* p, int value, size_t bytes)
{
    if (!p && bytes) return fail(VPS_ERR_INVALID, "null pointer");
    if (bytes) VPS_HIP(hipMemset(p, value, bytes));
    return VPS_OK;
}

extern "C" int vps_device_memcpy(void* dst, const void* src, size_t bytes, int to_device)
{
    if ((!dst || !src) && bytes) return fail(VPS_ERR_INVALID, "null pointer");
    if (bytes) VPS_HIP(hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    return VPS_OK;
}

extern "C" int vps_device_synchronize(int device)
{
    VPS_HIP(hipSetDevice(device));
    VPS_HIP(hipDeviceSynchronize());
    return VPS_OK;
}

extern "C" int vps_device_reset(int device)
{
    VPS_HIP(hipSetDevice(device));
    VPS_HIP(hipDeviceReset());
    return VPS_OK;
}

extern "C" int vps_device_name(int device, char* buf, size_t n)
{
    if (!buf || !n) return fail(VPS_ERR_INVALID, "null argument");
    hipDeviceProp_t prop;
    VPS_HIP(hipGetDeviceProperties(&prop, device));
    snprintf(buf, n, "%s (%s), %d compute units, %.1f GiB", prop.name, prop.gcnArchName, prop.multiProcessorCount,
             prop.totalGlobalMem / 1073741824.0);
    return VPS_OK;
}
