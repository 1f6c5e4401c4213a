// Shared host-side definitions of the SDPX GPU library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sdpx_gpu.h"

namespace sdpx {

// MPFR kind codes of the value layout (crates/arithmetic/src/lib.rs:333-340).
constexpr int KIND_NAN = 0;
constexpr int KIND_INF = 1;
constexpr int KIND_ZERO = 2;
constexpr int KIND_REGULAR = 3;
// Largest supported mantissa, in 64-bit limbs (1024-bit values).
constexpr int MAX_LIMBS = 16;
// Largest admitted per-operand exponent spread, in bits: the decline rule of
// crates/solver/src/algebra/dense/blas/rns_blas.rs:40,2198-2201.
constexpr int64_t MAX_SPREAD = 4096;

// An error carrying an ABI status code; caught at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

void hip_check(hipError_t e, const char *what);
#define HIP_CHECK(x) ::sdpx::hip_check((x), #x)

// Grow-only device buffers, one set per library (calls are serialised by the
// device lock).  Contents do not survive a later call.
struct Workspace {
    struct Buf {
        void *ptr = nullptr;
        size_t size = 0;
    };
    std::vector<Buf> bufs;
    void *get(int slot, size_t bytes);
    void release_all();
};

// Pinned host staging for small uploads on the library stream: a ring whose
// wrap first waits for the stream, so staged bytes are never overwritten
// before their copy has run.
struct PinnedRing {
    uint8_t *base = nullptr;
    size_t size = 0, top = 0;
};

// The library's device state: current device, its stream and workspace.
struct Device {
    int id = -1;
    hipStream_t stream = nullptr;
    Workspace ws;
    PinnedRing pin;
    int64_t *readback = nullptr;  // pinned, 16 words that kernels write for the host to read after a sync
    std::recursive_mutex mu;
};
Device &device();  // initialised on first use (device 0 unless selected)
void device_select(int id);
// Host-to-device copy of `bytes` from src (any host memory, free to reuse on
// return) into device memory dst, queued on the library stream without a
// wait.
void upload_async(void *dst, const void *src, size_t bytes);
// The device's pinned read-back words (16 int64), allocated on first use.
int64_t *readback_words();

// Value layout helpers: a value is 2N+2 u32 words (limbs, kind, exponent).
inline size_t value_words(int n) { return 2 * (size_t)n + 2; }
inline size_t value_bytes(int n) { return 8 * (size_t)n + 8; }

}  // namespace sdpx
