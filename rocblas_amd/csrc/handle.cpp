// handle.cpp -- rocblas_handle life cycle, stream / pointer mode, workspace and
// the device-memory size-query protocol.  Behaviour follows the reference's
// library/src/rocblas_auxiliary.cpp and library/src/handle.cpp (cited per
// function); the implementation is this library's own.
#include "internal.hpp"

#include "../../include/rocblas_amd_ext.h"
#include "logging.hpp"

#include <cstdlib>
#include <new>

namespace rbamd
{
rocblas_status hip_to_status(hipError_t e)
{
    switch(e)
    {
    case hipSuccess:
        return rocblas_status_success;
    case hipErrorMemoryAllocation:
    case hipErrorLaunchOutOfResources:
        return rocblas_status_memory_error;
    case hipErrorInvalidDevicePointer:
        return rocblas_status_invalid_pointer;
    case hipErrorInvalidDevice:
    case hipErrorInvalidResourceHandle:
        return rocblas_status_invalid_handle;
    case hipErrorInvalidValue:
        return rocblas_status_invalid_value;
    case hipErrorNoDevice:
    case hipErrorUnknown:
    default:
        return rocblas_status_internal_error;
    }
}

// Grow-on-demand rule of library/src/handle.cpp:376-407: a library-managed
// workspace that is too small is freed and reallocated at the requested size
// plus the default, and only while no lease is outstanding.
void* workspace_acquire(rocblas_handle h, size_t bytes)
{
    bytes = roundup_ws(bytes);
    if(h->ws_in_use)
        return nullptr;
    if(bytes > h->ws_size)
    {
        if(h->owner != WorkspaceOwner::library_managed)
            return nullptr;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if(h->stream && hipStreamIsCapturing(h->stream, &cs) == hipSuccess
           && cs != hipStreamCaptureStatusNone)
            return nullptr; // no allocation inside a captured region
        DeviceGuard g(h->device);
        if(h->ws)
        {
            if(hipFree(h->ws) != hipSuccess)
                return nullptr;
            h->ws      = nullptr;
            h->ws_size = 0;
        }
        size_t total = bytes + h->default_ws;
        if(hipMalloc(&h->ws, total) != hipSuccess)
        {
            h->ws = nullptr;
            return nullptr;
        }
        h->ws_size = total;
    }
    h->ws_in_use = true;
    return h->ws;
}

void workspace_release(rocblas_handle h)
{
    h->ws_in_use = false;
}

static rocblas_status free_workspace(rocblas_handle h)
{
    if(h->ws_in_use)
        return rocblas_status_internal_error;
    if(h->ws && h->owner != WorkspaceOwner::user_owned)
    {
        DeviceGuard g(h->device);
        hipError_t  e = hipFree(h->ws);
        if(e != hipSuccess)
            return hip_to_status(e);
    }
    h->ws      = nullptr;
    h->ws_size = 0;
    h->owner   = WorkspaceOwner::library_managed;
    return rocblas_status_success;
}
} // namespace rbamd

using namespace rbamd;

extern "C" {

// rocblas_auxiliary.cpp:188-204 and the handle constructor at
// library/src/handle.cpp:167-260: bind the current device, read
// ROCBLAS_DEVICE_MEMORY_SIZE, allocate the default workspace.
rocblas_status rocblas_create_handle(rocblas_handle* handle)
try
{
    if(!handle)
        return rocblas_status_invalid_handle;
    *handle = nullptr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if(e != hipSuccess)
        return hip_to_status(e);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if(e != hipSuccess)
        return hip_to_status(e);

    auto* h       = new _rocblas_handle;
    h->device     = dev;
    h->num_cus    = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    h->max_lds    = prop.sharedMemPerBlock;
    h->default_ws = kDefaultWorkspace;

    const char* env = std::getenv("ROCBLAS_DEVICE_MEMORY_SIZE");
    size_t      want = h->default_ws;
    if(env)
    {
        want = std::strtoull(env, nullptr, 0);
        if(want)
            h->owner = WorkspaceOwner::user_managed;
    }
    if(want)
    {
        want = roundup_ws(want);
        if(hipMalloc(&h->ws, want) != hipSuccess)
        {
            delete h;
            return rocblas_status_memory_error;
        }
        h->ws_size = want;
    }
    // completion counters of the chained solve and of the split-K GEMM, zero
    // between launches, and the host-visible error word their bounded waits
    // raise
    constexpr size_t flag_bytes
        = rbamd::kChainFlagsBytes + rbamd::kSplitCounters * sizeof(unsigned);
    if(hipMalloc(&h->chain_flags, flag_bytes) != hipSuccess
       || hipMemset(h->chain_flags, 0, flag_bytes) != hipSuccess
       || hipHostMalloc(reinterpret_cast<void**>(&h->async_error), sizeof(unsigned int),
                        hipHostMallocCoherent)
              != hipSuccess
       || hipDeviceSynchronize() != hipSuccess)
    {
        if(h->async_error)
            (void)hipHostFree(h->async_error);
        if(h->chain_flags)
            (void)hipFree(h->chain_flags);
        if(h->ws)
            (void)hipFree(h->ws);
        delete h;
        return rocblas_status_memory_error;
    }
    *h->async_error = 0u;
    rbamd::init_logging(h);
    *handle = h;
    return rocblas_status_success;
}
catch(const std::bad_alloc&)
{
    return rocblas_status_memory_error;
}
catch(...)
{
    return rocblas_status_internal_error;
}

// rocblas_auxiliary.cpp:211-227
rocblas_status rocblas_destroy_handle(rocblas_handle handle)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    handle->ws_in_use = false;
    (void)free_workspace(handle);
    if(handle->chain_flags)
    {
        (void)hipDeviceSynchronize();
        (void)hipFree(handle->chain_flags);
    }
    if(handle->async_error)
        (void)hipHostFree(handle->async_error);
    delete handle;
    return rocblas_status_success;
}

// rocblas_auxiliary.cpp:234-276: unchanged stream is a no-op; a non-null
// stream must be a valid stream (hipStreamQuery != invalid handle), checked
// only when neither stream is being captured.
rocblas_status rocblas_set_stream(rocblas_handle handle, hipStream_t stream)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(stream == handle->stream)
        return rocblas_status_success;
    hipStreamCaptureStatus new_cs = hipStreamCaptureStatusNone;
    if(stream)
    {
        if(hipStreamIsCapturing(stream, &new_cs) != hipSuccess)
            return rocblas_status_invalid_value;
    }
    hipStreamCaptureStatus old_cs = hipStreamCaptureStatusNone;
    if(handle->stream)
        (void)hipStreamIsCapturing(handle->stream, &old_cs);
    if(old_cs == hipStreamCaptureStatusNone && new_cs == hipStreamCaptureStatusNone && stream
       && hipStreamQuery(stream) == hipErrorInvalidHandle)
        return rocblas_status_invalid_value;
    handle->stream = stream;
    return rocblas_status_success;
}

// rocblas_auxiliary.cpp:282-296
rocblas_status rocblas_get_stream(rocblas_handle handle, hipStream_t* stream)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(!stream)
        return rocblas_status_invalid_pointer;
    *stream = handle->stream;
    return rocblas_status_success;
}

// rocblas_auxiliary.cpp:68-82 (the mode is stored as given)
rocblas_status rocblas_set_pointer_mode(rocblas_handle handle, rocblas_pointer_mode mode)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    handle->pointer_mode = mode;
    return rocblas_status_success;
}

// rocblas_auxiliary.cpp:48-63
rocblas_status rocblas_get_pointer_mode(rocblas_handle handle, rocblas_pointer_mode* mode)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(!mode)
        return rocblas_status_invalid_pointer;
    *mode = handle->pointer_mode;
    return rocblas_status_success;
}

// rocblas_auxiliary.cpp:783-810
const char* rocblas_status_to_string(rocblas_status status)
{
#define RBAMD_CASE(x) \
    case x:           \
        return #x
    switch(status)
    {
        RBAMD_CASE(rocblas_status_success);
        RBAMD_CASE(rocblas_status_invalid_handle);
        RBAMD_CASE(rocblas_status_not_implemented);
        RBAMD_CASE(rocblas_status_invalid_pointer);
        RBAMD_CASE(rocblas_status_invalid_size);
        RBAMD_CASE(rocblas_status_memory_error);
        RBAMD_CASE(rocblas_status_internal_error);
        RBAMD_CASE(rocblas_status_perf_degraded);
        RBAMD_CASE(rocblas_status_size_query_mismatch);
        RBAMD_CASE(rocblas_status_size_increased);
        RBAMD_CASE(rocblas_status_size_unchanged);
        RBAMD_CASE(rocblas_status_invalid_value);
        RBAMD_CASE(rocblas_status_continue);
        RBAMD_CASE(rocblas_status_check_numerics_fail);
        RBAMD_CASE(rocblas_status_excluded_from_build);
        RBAMD_CASE(rocblas_status_arch_mismatch);
    }
#undef RBAMD_CASE
    return "<undefined rocblas_status value>";
}

// handle.cpp:449-462
rocblas_status rocblas_start_device_memory_size_query(rocblas_handle handle)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(handle->size_query)
        return rocblas_status_size_query_mismatch;
    handle->size_query = true;
    handle->query_size = 0;
    return rocblas_status_success;
}

// handle.cpp:468-483
rocblas_status rocblas_stop_device_memory_size_query(rocblas_handle handle, size_t* size)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(!handle->size_query)
        return rocblas_status_size_query_mismatch;
    if(!size)
        return rocblas_status_invalid_pointer;
    *size              = handle->query_size;
    handle->size_query = false;
    return rocblas_status_success;
}

// handle.cpp:654-657
bool rocblas_is_device_memory_size_query(rocblas_handle handle)
{
    return handle && handle->size_query;
}

// handle.cpp:489-500
rocblas_status rocblas_get_device_memory_size(rocblas_handle handle, size_t* size)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(!size)
        return rocblas_status_invalid_pointer;
    *size = handle->ws_size;
    return rocblas_status_success;
}

// handle.cpp:541-590: free what the library owns; size 0 returns to
// library management with deferred allocation; otherwise allocate the rounded
// size and mark it user-managed.
rocblas_status rocblas_set_device_memory_size(rocblas_handle handle, size_t size)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    DeviceGuard    g(handle->device);
    rocblas_status st = free_workspace(handle);
    if(st != rocblas_status_success)
        return st;
    if(!size)
        return rocblas_status_success;
    size           = roundup_ws(size);
    hipError_t e   = hipMalloc(&handle->ws, size);
    if(e != hipSuccess)
    {
        handle->ws = nullptr;
        return hip_to_status(e);
    }
    handle->ws_size = size;
    handle->owner   = WorkspaceOwner::user_managed;
    return rocblas_status_success;
}

// handle.cpp:596-624
rocblas_status rocblas_set_workspace(rocblas_handle handle, void* addr, size_t size)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    DeviceGuard    g(handle->device);
    rocblas_status st = free_workspace(handle);
    if(st != rocblas_status_success)
        return st;
    if(size && addr)
    {
        handle->owner   = WorkspaceOwner::user_owned;
        handle->ws_size = size;
        handle->ws      = addr;
    }
    return rocblas_status_success;
}

// handle.cpp:630-638
bool rocblas_is_managing_device_memory(rocblas_handle handle)
{
    return handle && handle->owner == WorkspaceOwner::library_managed;
}

// handle.cpp:643-650
bool rocblas_is_user_managing_device_memory(rocblas_handle handle)
{
    return handle && handle->owner == WorkspaceOwner::user_managed;
}

// No counterpart in the reference.  Completes the handle's stream (unless it is
// being captured) and reports whether a kernel of an earlier call gave up a
// bounded wait, clearing the report.
rocblas_status rocblas_amd_get_async_status(rocblas_handle handle)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if(handle->stream)
        (void)hipStreamIsCapturing(handle->stream, &cs);
    if(cs == hipStreamCaptureStatusNone)
    {
        rbamd::DeviceGuard dg(handle->device);
        hipError_t e = hipStreamSynchronize(handle->stream);
        if(e != hipSuccess)
            return rbamd::hip_to_status(e);
    }
    return rbamd::take_async_error(handle);
}

rocblas_status rocblas_amd_set_chain_test_timeout(rocblas_handle handle, int enable)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    handle->chain_debug_timeout = enable != 0;
    return rocblas_status_success;
}
} // extern "C"
