// GPU self-test of the wave-distributed arithmetic (wavemp.h) against the
// register routines of mpdev.h on the same device: each case runs on one
// wave, once in register form (lane 0 stores) and once spread over the wave.
// Test infrastructure only; built into tests/native/libwavemp_selftest.so.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../sdpx.rs_amd/csrc/wavemp.h"

using namespace sdpx;

namespace {

template <int W>
constexpr int slots() { return 2 * W + 20; }

// Serial restatement of an exact window's rounding: signed 64-bit slot
// totals t[0..S) (weight 2^(32 s) at scale base), carried word by word,
// negated when the total is negative, rounded once by round_words.
template <int W, int S>
__device__ void window_serial(int64_t (&t)[S], int64_t base, mpd::Val<W> &out) {
    uint32_t w[S];
    int64_t carry = 0;
    for (int s = 0; s < S; s++) {
        const int64_t v = t[s] + carry;
        w[s] = (uint32_t)(uint64_t)v;
        carry = v >> 32;
    }
    const bool neg = carry < 0;
    if (neg) {
        uint64_t c = 1;
        for (int s = 0; s < S; s++) {
            const uint64_t v = (uint64_t)(~w[s]) + c;
            w[s] = (uint32_t)v;
            c = v >> 32;
        }
    }
    mpd::round_words<W, S>(neg, w, base, out);
}

// op: 0 a + b, 1 a - b, 2 a * b, 3 1 / a, 4 sign of a - b (word 0 of the
// output, +1 / 0 / -1), 5 rounding of the window T, 6 the window T plus
// (a * b) << off, then its rounding, 7 RN(c - a * b) with c the record at
// the start of the case's row of T.
template <int W>
__global__ __launch_bounds__(64) void wmp_case(int op, const uint32_t *A, const uint32_t *B, const int64_t *T,
                                               const int64_t *bo, int n, uint32_t *S, uint32_t *O) {
    constexpr int SL = slots<W>();
    const int c = blockIdx.x;
    if (c >= n) return;
    const int k = threadIdx.x;
    const uint32_t *a = A + (size_t)c * (W + 2), *b = B + (size_t)c * (W + 2);
    uint32_t *so = S + (size_t)c * (W + 2), *wo = O + (size_t)c * (W + 2);
    mpd::Val<W> va, vb, vr;
    mpd::load(a, va);
    mpd::load(b, vb);
    const wmp::WV<W> wa = wmp::load<W>(a), wb = wmp::load<W>(b);
    wmp::WV<W> wr = wmp::zero<W>();
    const int64_t base = bo[2 * c], off = bo[2 * c + 1];
    const int64_t tl = k < SL ? T[(size_t)c * 64 + k] : 0;
    if (op == 0 || op == 1) {
        mpd::add_signed(va, vb, op == 1, vr);
        wr = wmp::add(wa, wb, op == 1);
    } else if (op == 2) {
        mpd::mul(va, vb, vr);
        wr = wmp::mul(wa, wb);
    } else if (op == 3) {
        mpd::recip(va, vr);
        wr = wmp::recip(wa);
    } else if (op == 4) {
        mpd::Val<W> d;
        mpd::add_signed(va, vb, true, d);
        mpd::set_zero(vr);
        vr.m[0] = mpd::is_zero(d) ? 0u : (d.kind < 0 ? 0xFFFFFFFFu : 1u);
        const int s = wmp::compare(wa, wb);
        wr.w = k == 0 ? (uint32_t)s : 0u;
    } else if (op == 7) {
        const uint32_t *cr = (const uint32_t *)(T + (size_t)c * 64);
        mpd::Val<W> vc;
        mpd::load(cr, vc);
        mpd::fms(vc, va, vb, vr);
        wr = wmp::fms(wmp::load<W>(cr), wa, wb);
    } else {
        int64_t t[SL];
        for (int s = 0; s < SL; s++) t[s] = T[(size_t)c * 64 + s];
        int64_t tw = tl;
        if (op == 6 && !mpd::is_zero(va) && !mpd::is_zero(vb)) {
            uint32_t p[2 * W + 1];
            for (int i = 0; i <= 2 * W; i++) p[i] = 0;
            for (int i = 0; i < W; i++) {
                uint64_t carry = 0;
                for (int j = 0; j < W; j++) {
                    const uint64_t v = (uint64_t)va.m[i] * vb.m[j] + p[i + j] + carry;
                    p[i + j] = (uint32_t)v;
                    carry = v >> 32;
                }
                p[i + W] = (uint32_t)carry;
            }
            const int q = (int)(off >> 5), r = (int)(off & 31);
            const bool neg = (va.kind < 0) != (vb.kind < 0);
            for (int i = 0; i <= 2 * W; i++) {
                const uint32_t hi = i < 2 * W ? p[i] : 0u, lo = i ? p[i - 1] : 0u;
                const uint32_t word = r ? (uint32_t)((((uint64_t)hi << 32) | lo) >> (32 - r)) : hi;
                if (q + i < SL) t[q + i] += neg ? -(int64_t)word : (int64_t)word;
            }
            tw = wmp::add_product<W, SL>(tw, wa, wb, off, false);
        }
        window_serial<W, SL>(t, base, vr);
        wr = wmp::round_window<W, SL>(tw, base);
    }
    if (k == 0) mpd::store(so, vr);
    wmp::store(wo, wr);
}

}  // namespace

extern "C" int wmp_selftest(int limbs, int op, const uint32_t *a, const uint32_t *b, const int64_t *T,
                            const int64_t *base_off, int n, uint32_t *scalar_out, uint32_t *wave_out) {
    const int W = 2 * limbs;
    if (n <= 0 || !(W == 4 || W == 8 || W == 16)) return -1;
    const size_t rb = (size_t)n * (W + 2) * 4, tb = (size_t)n * 64 * 8, ob = (size_t)n * 16;
    uint32_t *da = nullptr, *db = nullptr, *ds = nullptr, *dw = nullptr;
    int64_t *dt = nullptr, *dbo = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    step(hipMalloc(&da, rb));
    step(hipMalloc(&db, rb));
    step(hipMalloc(&ds, rb));
    step(hipMalloc(&dw, rb));
    step(hipMalloc(&dt, tb));
    step(hipMalloc(&dbo, ob));
    if (e == hipSuccess) {
        step(hipMemcpy(da, a, rb, hipMemcpyHostToDevice));
        step(hipMemcpy(db, b, rb, hipMemcpyHostToDevice));
        step(hipMemcpy(dt, T, tb, hipMemcpyHostToDevice));
        step(hipMemcpy(dbo, base_off, ob, hipMemcpyHostToDevice));
        step(hipMemset(ds, 0, rb));
        step(hipMemset(dw, 0, rb));
    }
    if (e == hipSuccess) {
        if (W == 4)
            hipLaunchKernelGGL(wmp_case<4>, dim3(n), dim3(64), 0, 0, op, da, db, dt, dbo, n, ds, dw);
        else if (W == 8)
            hipLaunchKernelGGL(wmp_case<8>, dim3(n), dim3(64), 0, 0, op, da, db, dt, dbo, n, ds, dw);
        else
            hipLaunchKernelGGL(wmp_case<16>, dim3(n), dim3(64), 0, 0, op, da, db, dt, dbo, n, ds, dw);
        step(hipGetLastError());
        step(hipDeviceSynchronize());
    }
    if (e == hipSuccess) {
        step(hipMemcpy(scalar_out, ds, rb, hipMemcpyDeviceToHost));
        step(hipMemcpy(wave_out, dw, rb, hipMemcpyDeviceToHost));
    }
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(ds);
    (void)hipFree(dw);
    (void)hipFree(dt);
    (void)hipFree(dbo);
    return (int)e;
}
