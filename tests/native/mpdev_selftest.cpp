// Host instantiation of the GPU panel arithmetic (mpdev.h) against the host
// library's arithmetic (mpx.h), bit for bit, on random values.
// Build: hipcc -x hip --offload-host-only -O2 -std=c++17 (see test_mpdev.py).
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../sdpx.rs_amd/csrc/mpdev.h"
#include "../../sdpx.rs_amd/csrc/mpx.h"

using namespace sdpx;

template <int N>
void random_value(std::mt19937_64 &rng, uint8_t *rec, int spread) {
    uint64_t *l = (uint64_t *)rec;
    for (int i = 0; i < N; i++) l[i] = rng();
    l[N - 1] |= 1ULL << 63;
    if (rng() % 7 == 0) l[0] &= ~0xFFFFULL;  // some trailing zeros
    int32_t kind = (rng() & 1) ? -3 : 3;
    if (rng() % 23 == 0) {
        std::memset(rec, 0, 8 * N);
        kind = 2;
    }
    *(int32_t *)(rec + 8 * N) = kind;
    *(int32_t *)(rec + 8 * N + 4) = kind == 2 ? 0 : (int32_t)(rng() % (2 * spread + 1)) - spread;
}

template <int N>
int run(int count, int spread, uint64_t seed) {
    constexpr int W = 2 * N;
    const size_t vb = 8 * N + 8;
    std::mt19937_64 rng(seed);
    std::vector<uint8_t> a(vb), b(vb), r1(vb), r2(vb);
    int bad = 0;
    for (int t = 0; t < count; t++) {
        random_value<N>(rng, a.data(), spread);
        random_value<N>(rng, b.data(), rng() % 4 == 0 ? 2 * 32 * N + 40 : spread);
        mpd::Val<W> va, vb2, vr;
        mpd::load((const uint32_t *)a.data(), va);
        mpd::load((const uint32_t *)b.data(), vb2);
        mpx::View xa{N, a.data()}, xb{N, b.data()};
        // mul
        mpd::mul(va, vb2, vr);
        mpd::store((uint32_t *)r1.data(), vr);
        mpx::mul(xa, xb, r2.data());
        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("mul mismatch %d\n", t);
        // sub / add
        mpd::add_signed(va, vb2, true, vr);
        mpd::store((uint32_t *)r1.data(), vr);
        mpx::sub(xa, xb, r2.data());
        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("sub mismatch %d\n", t);
        mpd::add_signed(va, vb2, false, vr);
        mpd::store((uint32_t *)r1.data(), vr);
        mpx::add(xa, xb, r2.data());
        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("add mismatch %d\n", t);
        // recip
        if (!mpd::is_zero(va)) {
            mpd::recip(va, vr);
            mpd::store((uint32_t *)r1.data(), vr);
            mpx::recip(xa, r2.data());
            if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("recip mismatch %d\n", t);
        }
    }
    // exact dots with the slot window, against mpx's accumulator
    constexpr int SLOTS = 4 * N + 24;
    mpd::Val<W> vr;
    std::vector<int64_t> slots(SLOTS * 3);
    for (int t = 0; t < count / 4; t++) {
        const int k = 1 + (int)(rng() % 64);
        std::vector<uint8_t> xs(k * vb), ys(k * vb), c(vb);
        for (int i = 0; i < k; i++) {
            random_value<N>(rng, &xs[i * vb], 40);
            random_value<N>(rng, &ys[i * vb], 40);
        }
        random_value<N>(rng, c.data(), 60);
        for (int with_c = 0; with_c < 2; with_c++) {
            mpx::DotAcc acc;
            acc.dot(xs.data(), vb, ys.data(), vb, k, N, r2.data(), with_c ? c.data() : nullptr);
            // window base: min scale of the terms
            int64_t lo = INT64_MAX, hi = INT64_MIN;
            for (int i = 0; i < k; i++) {
                mpx::View x{N, &xs[i * vb]}, y{N, &ys[i * vb]};
                if (x.zero() || y.zero()) continue;
                int64_t s = (int64_t)x.exponent() + y.exponent() - 2 * 64 * N;
                lo = std::min(lo, s);
                hi = std::max(hi, s);
            }
            mpx::View cv{N, c.data()};
            if (with_c && !cv.zero()) {
                lo = std::min<int64_t>(lo, (int64_t)cv.exponent() - 64 * N);
                hi = std::max<int64_t>(hi, (int64_t)cv.exponent() - 2 * 64 * N);
            }
            if (lo == INT64_MAX) continue;
            if ((hi - lo) / 32 + 2 * W + 4 > SLOTS) continue;  // outside the window: host path
            mpd::DotWindow<W, SLOTS> dw{slots.data(), 3, lo};
            dw.clear();
            for (int i = 0; i < k; i++) {
                mpd::Val<W> x, y;
                mpd::load((const uint32_t *)&xs[i * vb], x);
                mpd::load((const uint32_t *)&ys[i * vb], y);
                if (mpd::is_zero(x) || mpd::is_zero(y)) continue;
                dw.add_product(x, y, with_c != 0);
            }
            if (with_c) {
                mpd::Val<W> cvv;
                mpd::load((const uint32_t *)c.data(), cvv);
                if (!mpd::is_zero(cvv)) dw.add_value(cvv, false);
            }
            dw.finish(vr);
            mpd::store((uint32_t *)r1.data(), vr);
            if (std::memcmp(r1.data(), r2.data(), vb)) {
                bad++;
                std::printf("dot mismatch %d k=%d c=%d\n", t, k, with_c);
                if (std::getenv("DUMP") && bad == 1) {
                    FILE *f = fopen("/tmp/dotcase.bin", "wb");
                    int32_t hdr[2] = {N, k};
                    fwrite(hdr, 4, 2, f);
                    fwrite(xs.data(), 1, k * vb, f);
                    fwrite(ys.data(), 1, k * vb, f);
                    fwrite(c.data(), 1, vb, f);
                    fwrite(r1.data(), 1, vb, f);
                    fwrite(r2.data(), 1, vb, f);
                    fclose(f);
                }
            }
        }
    }
    return bad;
}

// Reciprocals of the mantissas at the edges of the quotient-digit estimate:
// powers of two, all ones, a single low bit, a high word followed by zeros
// or ones, and the digit patterns that force the estimate's corrections.
template <int N>
int recip_edges() {
    constexpr int W = 2 * N;
    const size_t vb = 8 * N + 8;
    std::vector<uint8_t> a(vb), r1(vb), r2(vb);
    int bad = 0;
    const uint32_t pats[][2] = {{0x80000000u, 0}, {0xFFFFFFFFu, 0xFFFFFFFFu}, {0x80000000u, 1},
                                {0xFFFFFFFFu, 0}, {0x80000000u, 0xFFFFFFFFu}, {0xC0000000u, 0},
                                {0x80000001u, 0x80000000u}, {0xFFFFFFFEu, 0xFFFFFFFFu}};
    for (auto &pt : pats)
        for (int low = 0; low < 3; low++)
            for (int sgn = 0; sgn < 2; sgn++) {
                uint32_t *m = (uint32_t *)a.data();
                for (int i = 0; i < W; i++) m[i] = pt[1];
                m[W - 1] = pt[0];
                if (low == 1) m[0] = 1;
                if (low == 2 && W > 1) m[W - 2] = 0;
                *(int32_t *)(a.data() + 8 * N) = sgn ? -3 : 3;
                *(int32_t *)(a.data() + 8 * N + 4) = 17 - 40 * low;
                mpd::Val<W> va, vr;
                mpd::load((const uint32_t *)a.data(), va);
                mpd::recip(va, vr);
                mpd::store((uint32_t *)r1.data(), vr);
                mpx::recip(mpx::View{N, a.data()}, r2.data());
                if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("recip edge mismatch N=%d\n", N);
            }
    return bad;
}

// Sums, differences and products at the edges of the rounding paths:
// exponent gaps around each word boundary, 64 and P + 2, mantissas that
// cancel to any depth, exact ties (a power of two plus half an ulp),
// all-ones mantissas whose rounding carries out, against mpx.
template <int N>
int add_mul_edges(uint64_t seed) {
    constexpr int W = 2 * N;
    constexpr int P = 64 * N;
    const size_t vb = 8 * N + 8;
    std::mt19937_64 rng(seed);
    std::vector<uint8_t> a(vb), b(vb), r1(vb), r2(vb);
    int bad = 0;
    std::vector<int> gaps = {0, 1, 2, 3, 31, 32, 33, 63, 64, 65, 66, 95, 96, 97, P - 2, P - 1, P, P + 1, P + 2, P + 3};
    auto setm = [&](uint8_t *rec, int pat, const uint8_t *like) {
        uint32_t *m = (uint32_t *)rec;
        for (int i = 0; i < W; i++) {
            switch (pat) {
                case 0: m[i] = 0; break;                                    // power of two
                case 1: m[i] = 0xFFFFFFFFu; break;                          // all ones
                case 2: m[i] = i == 0 ? 1u : 0u; break;                      // power of two + 1 ulp
                case 3: m[i] = (uint32_t)rng(); break;                       // random
                case 4: m[i] = ((const uint32_t *)like)[i]; break;           // same as a
                case 5: m[i] = ((const uint32_t *)like)[i] ^ (i < W / 2 ? (uint32_t)rng() : 0u); break;  // low half differs
                case 6: m[i] = ((const uint32_t *)like)[i] - (i == 0 ? 1u : 0u); break;  // one ulp below a (a's low word non-zero)
                default: m[i] = i < W - 1 ? 0u : 0x40000000u; break;          // top bit clear until fixed below
            }
        }
        m[W - 1] |= 0x80000000u;
    };
    for (int gap : gaps)
        for (int pa = 0; pa < 4; pa++)
            for (int pb = 0; pb < 7; pb++)
                for (int sg = 0; sg < 4; sg++) {
                    setm(a.data(), pa, nullptr);
                    if (pa == 3) ((uint32_t *)a.data())[0] |= 1u;
                    setm(b.data(), pb, a.data());
                    const int ea = (int)(rng() % 200) - 100;
                    *(int32_t *)(a.data() + 8 * N) = (sg & 1) ? -3 : 3;
                    *(int32_t *)(a.data() + 8 * N + 4) = ea;
                    *(int32_t *)(b.data() + 8 * N) = (sg & 2) ? -3 : 3;
                    *(int32_t *)(b.data() + 8 * N + 4) = (pb >= 4 && gap > 3) ? ea : ea - gap;
                    for (int swap = 0; swap < 2; swap++) {
                        const uint8_t *pa_ = swap ? b.data() : a.data(), *pb_ = swap ? a.data() : b.data();
                        mpd::Val<W> va, vb2, vr;
                        mpd::load((const uint32_t *)pa_, va);
                        mpd::load((const uint32_t *)pb_, vb2);
                        mpx::View xa{N, pa_}, xb{N, pb_};
                        mpd::add_signed(va, vb2, false, vr);
                        mpd::store((uint32_t *)r1.data(), vr);
                        mpx::add(xa, xb, r2.data());
                        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("add edge mismatch N=%d gap=%d\n", N, gap);
                        mpd::add_signed(va, vb2, true, vr);
                        mpd::store((uint32_t *)r1.data(), vr);
                        mpx::sub(xa, xb, r2.data());
                        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("sub edge mismatch N=%d gap=%d\n", N, gap);
                        mpd::mul(va, vb2, vr);
                        mpd::store((uint32_t *)r1.data(), vr);
                        mpx::mul(xa, xb, r2.data());
                        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("mul edge mismatch N=%d\n", N);
                    }
                }
    return bad;
}

// Division and square root against the C oracle (liboracle.so, path in
// argv[2]): both are correctly rounded, so the bits must agree.
typedef void (*ora_div_fn)(int, const uint8_t *, const uint8_t *, uint8_t *);
typedef void (*ora_sqrt_fn)(int, const uint8_t *, uint8_t *);
template <int N>
int div_sqrt_vs_oracle(ora_div_fn odiv, ora_sqrt_fn osqrt, int count, uint64_t seed) {
    constexpr int W = 2 * N;
    const size_t vb = 8 * N + 8;
    std::mt19937_64 rng(seed);
    std::vector<uint8_t> a(vb), b(vb), r1(vb), r2(vb);
    int bad = 0;
    for (int t = 0; t < count; t++) {
        random_value<N>(rng, a.data(), 300);
        random_value<N>(rng, b.data(), 300);
        if (t < 4) {  // power of two, all ones, one low bit
            uint32_t *m = (uint32_t *)a.data();
            for (int i = 0; i < W; i++) m[i] = t == 1 ? 0xFFFFFFFFu : 0u;
            m[W - 1] |= 0x80000000u;
            if (t == 2) m[0] |= 1u;
            *(int32_t *)(a.data() + 8 * N) = 3;
            *(int32_t *)(a.data() + 8 * N + 4) = t == 3 ? 7 : 0;
        }
        mpd::Val<W> va, vb2, vr;
        mpd::load((const uint32_t *)a.data(), va);
        mpd::load((const uint32_t *)b.data(), vb2);
        if (!mpd::is_zero(vb2)) {
            mpd::div(va, vb2, vr);
            mpd::store((uint32_t *)r1.data(), vr);
            odiv(N, a.data(), b.data(), r2.data());
            if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("div mismatch N=%d %d\n", N, t);
        }
        if (!mpd::is_zero(va)) {
            va.kind = mpd::KIND_REG;  // positive radicand
            *(int32_t *)(a.data() + 8 * N) = 3;
            mpd::sqrt(va, vr);
            mpd::store((uint32_t *)r1.data(), vr);
            osqrt(N, a.data(), r2.data());
            if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("sqrt mismatch N=%d %d\n", N, t);
        }
    }
    return bad;
}

typedef void (*ora_fms_fn)(int, const uint8_t *, const uint8_t *, const uint8_t *, uint8_t *);

// RN(v - a b) against the oracle: random operands, v = RN(a b) (deep
// cancellation, also exact zero), v far above and far below the product,
// and zero operands.
template <int N>
int fms_vs_oracle(ora_fms_fn ofms, int count, uint64_t seed) {
    constexpr int W = 2 * N;
    const size_t vb = 8 * N + 8;
    std::mt19937_64 rng(seed);
    std::vector<uint8_t> v(vb), a(vb), b(vb), r1(vb), r2(vb);
    int bad = 0;
    for (int t = 0; t < count; t++) {
        random_value<N>(rng, v.data(), 300);
        random_value<N>(rng, a.data(), 300);
        random_value<N>(rng, b.data(), 300);
        mpd::Val<W> va, vb2, vv, vr;
        mpd::load((const uint32_t *)a.data(), va);
        mpd::load((const uint32_t *)b.data(), vb2);
        mpd::load((const uint32_t *)v.data(), vv);
        const int mode = t % 6;
        if (mode == 1 || mode == 2) {  // v = RN(a b), or it nudged by one unit
            if (!mpd::is_zero(va) && !mpd::is_zero(vb2)) {
                mpd::mul(va, vb2, vv);
                if (mode == 2) vv.m[0] ^= 1u;
            }
        } else if (mode == 3 && mpd::is_regular(vv) && mpd::is_regular(va) && mpd::is_regular(vb2)) {
            // v far below or far above the product, up to the buffer's width
            const int64_t pe = (int64_t)va.e + vb2.e;
            const int64_t off = (int64_t)(rng() % (4 * 64 * N + 40)) - (2 * 64 * N + 20);
            const int64_t e = pe + off;
            if (e > -100000 && e < 100000) vv.e = (int32_t)e;
        }
        if (mode == 4 && mpd::is_regular(vv)) {
            // a b exactly halfway between two W-word values (a = 3 * 2^k,
            // b's mantissa ending in binary 10) and v far below: only the
            // sign of v decides the rounding
            for (int i = 0; i < W; i++) va.m[i] = 0;
            va.m[W - 1] = 0xC0000000u;
            va.kind = mpd::KIND_REG;
            vb2.kind = mpd::KIND_REG;
            for (int i = 0; i < W; i++) vb2.m[i] = (uint32_t)rng();
            vb2.m[W - 1] |= 0x80000000u;
            vb2.m[0] = (vb2.m[0] & ~3u) | 2u;
            va.e = 5;
            vb2.e = -3;
            vv.e = 2 - 32 * W - (int32_t)(rng() % (96 * W));
            mpd::store((uint32_t *)a.data(), va);
            mpd::store((uint32_t *)b.data(), vb2);
        }
        mpd::store((uint32_t *)v.data(), vv);
        mpd::fms(vv, va, vb2, vr);
        mpd::store((uint32_t *)r1.data(), vr);
        ofms(N, v.data(), a.data(), b.data(), r2.data());
        if (std::memcmp(r1.data(), r2.data(), vb)) bad++, std::printf("fms mismatch N=%d %d mode %d\n", N, t, mode);
    }
    return bad;
}

int main(int argc, char **argv) {
    int count = argc > 1 ? std::atoi(argv[1]) : 20000;
    int bad = run<1>(count, 200, 1) + run<2>(count, 200, 2) + run<4>(count, 300, 3) +
              run<8>(count, 600, 4);
    bad += recip_edges<1>() + recip_edges<2>() + recip_edges<4>() + recip_edges<8>() + recip_edges<16>();
    bad += add_mul_edges<1>(31) + add_mul_edges<2>(32) + add_mul_edges<4>(33) + add_mul_edges<8>(34);
    if (argc > 2) {
        void *h = dlopen(argv[2], RTLD_NOW);
        if (!h) {
            std::printf("cannot load %s\n", argv[2]);
            return 2;
        }
        auto odiv = (ora_div_fn)dlsym(h, "oracle_div");
        auto osqrt = (ora_sqrt_fn)dlsym(h, "oracle_sqrt");
        auto ofms = (ora_fms_fn)dlsym(h, "oracle_fms");
        if (!odiv || !osqrt || !ofms) {
            std::printf("oracle symbols missing\n");
            return 2;
        }
        bad += fms_vs_oracle<1>(ofms, count / 4, 21) + fms_vs_oracle<2>(ofms, count / 4, 22) +
               fms_vs_oracle<4>(ofms, count / 4, 23) + fms_vs_oracle<8>(ofms, count / 4, 24);
        bad += div_sqrt_vs_oracle<1>(odiv, osqrt, count / 4, 11) + div_sqrt_vs_oracle<2>(odiv, osqrt, count / 4, 12) +
               div_sqrt_vs_oracle<4>(odiv, osqrt, count / 4, 13) + div_sqrt_vs_oracle<8>(odiv, osqrt, count / 4, 14);
    }
    std::printf("mismatches %d\n", bad);
    return bad ? 1 : 0;
}
