// Device selection, stream, workspace and HIP error handling.
#include <cstring>

#include "common.h"
#include "rns_gemm.h"

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

void upload_async(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    Device &d = device();
    constexpr size_t RING = 4u << 20;
    if (bytes > RING / 4) {  // large: a plain copy, waited for
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return;
    }
    if (!d.pin.base) {
        HIP_CHECK(hipHostMalloc((void **)&d.pin.base, RING, hipHostMallocDefault));
        d.pin.size = RING;
        d.pin.top = 0;
    }
    size_t off = (d.pin.top + 255) & ~size_t(255);
    if (off + bytes > d.pin.size) {  // wrap: every earlier staged copy has run
        HIP_CHECK(hipStreamSynchronize(d.stream));
        off = 0;
    }
    std::memcpy(d.pin.base + off, src, bytes);
    HIP_CHECK(hipMemcpyAsync(dst, d.pin.base + off, bytes, hipMemcpyHostToDevice, d.stream));
    d.pin.top = off + bytes;
}

int64_t *readback_words() {
    Device &d = device();
    // fine-grained (coherent): kernel writes reach the host without a cache flush
    if (!d.readback) HIP_CHECK(hipHostMalloc((void **)&d.readback, 16 * sizeof(int64_t), hipHostMallocCoherent));
    return d.readback;
}

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
    // set-up of the allocation-free exact product: this device's scratch address
    rns_noalloc_setup();
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
        if (g_device.stream) (void)hipStreamSynchronize(g_device.stream);
        if (g_device.pin.base) (void)hipHostFree(g_device.pin.base);
        g_device.pin = PinnedRing{};
        if (g_device.readback) (void)hipHostFree(g_device.readback);
        g_device.readback = nullptr;
        if (g_device.stream) (void)hipStreamDestroy(g_device.stream);
        g_device.stream = nullptr;
    }
    g_device.id = -1;
    init_device(id);
}

}  // namespace sdpx
