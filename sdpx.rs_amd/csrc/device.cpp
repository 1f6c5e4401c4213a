// Device selection, stream, workspace and HIP error handling.
#include <cstring>

#include "common.h"

namespace sdpx {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        throw Error(SDPX_GPU_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    }
}

void *Workspace::get(int slot, size_t bytes) {
    if ((int)bufs.size() <= slot) bufs.resize(slot + 1);
    Buf &b = bufs[slot];
    if (bytes == 0) bytes = 256;
    if (b.size < bytes) {
        if (b.ptr) HIP_CHECK(hipFree(b.ptr));
        b.ptr = nullptr;
        b.size = 0;
        // round up to limit regrowth
        size_t want = bytes + bytes / 8;
        want = (want + 4095) & ~size_t(4095);
        HIP_CHECK(hipMalloc(&b.ptr, want));
        b.size = want;
    }
    return b.ptr;
}

void Workspace::release_all() {
    for (auto &b : bufs) {
        if (b.ptr) (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.size = 0;
    }
}

static Device g_device;

static void init_device(int id) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw Error(SDPX_GPU_DEVICE_ERROR, "no HIP device present");
    if (id < 0 || id >= count) throw Error(SDPX_GPU_INVALID, "device index out of range");
    HIP_CHECK(hipSetDevice(id));
    if (g_device.stream == nullptr || g_device.id != id) {
        hipStream_t s;
        // default priority only
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        g_device.stream = s;
    }
    g_device.id = id;
}

Device &device() {
    if (g_device.id < 0) {
        std::lock_guard<std::recursive_mutex> g(g_device.mu);
        if (g_device.id < 0) init_device(0);
    }
    HIP_CHECK(hipSetDevice(g_device.id));
    return g_device;
}

void device_select(int id) {
    std::lock_guard<std::recursive_mutex> g(g_device.mu);
    if (g_device.id == id) return;
    if (g_device.id >= 0) {
        g_device.ws.release_all();
        if (g_device.stream) (void)hipStreamDestroy(g_device.stream);
        g_device.stream = nullptr;
    }
    g_device.id = -1;
    init_device(id);
}

}  // namespace sdpx
