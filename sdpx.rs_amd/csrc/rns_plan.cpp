#include "rns_plan.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <mutex>

namespace sdpx {
namespace {

bool is_prime(uint32_t n) {
    if (n < 2) return false;
    if (n % 2 == 0) return n == 2;
    for (uint32_t d = 3; d * d <= n; d += 2)
        if (n % d == 0) return false;
    return true;
}

struct Table {
    std::vector<uint32_t> q;
    int single = 0;
    Table() {
        // One-digit: the largest power <= 256 of every prime <= 256 with a
        // value >= 17 (smaller moduli carry under 4 bits per GEMM).
        std::vector<uint32_t> one;
        for (uint32_t p = 2; p <= 256; p++) {
            if (!is_prime(p)) continue;
            uint32_t v = p;
            while (v * p <= 256) v *= p;
            if (v >= 17) one.push_back(v);
        }
        std::sort(one.begin(), one.end(), std::greater<uint32_t>());
        q = one;
        single = (int)one.size();
        // Two-digit: primes in (256, 65279], descending.
        for (uint32_t p = 65279; p > 256; p--)
            if (is_prime(p)) q.push_back(p);
    }
};

const Table &table() {
    static Table t;
    return t;
}

// Little-endian base-2^32 big integer helpers (host only, plan setup).
void mul_small(std::vector<uint32_t> &a, uint32_t b) {
    uint64_t carry = 0;
    for (auto &x : a) {
        uint64_t t = (uint64_t)x * b + carry;
        x = (uint32_t)t;
        carry = t >> 32;
    }
    if (carry) a.push_back((uint32_t)carry);
}
uint32_t div_small(const std::vector<uint32_t> &a, uint32_t b, std::vector<uint32_t> &out) {
    out.assign(a.size(), 0);
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) {
        uint64_t cur = (rem << 32) | a[i];
        out[i] = (uint32_t)(cur / b);
        rem = cur % b;
    }
    return (uint32_t)rem;
}
uint32_t mod_small(const std::vector<uint32_t> &a, uint32_t b) {
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) rem = ((rem << 32) | a[i]) % b;
    return (uint32_t)rem;
}
uint32_t inverse(uint32_t a, uint32_t p) {
    int64_t t = 0, nt = 1, r = p, nr = a % p;
    while (nr) {
        int64_t qq = r / nr;
        int64_t tmp = t - qq * nt;
        t = nt;
        nt = tmp;
        tmp = r - qq * nr;
        r = nr;
        nr = tmp;
    }
    return (uint32_t)(t < 0 ? t + p : t);
}
uint32_t gcd(uint32_t a, uint32_t b) {
    while (b) {
        uint32_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}
// 24-bit limbs of a base-2^32 integer, exactly `limbs` of them.
std::vector<double> limbs24(const std::vector<uint32_t> &a, int limbs) {
    std::vector<double> out(limbs, 0.0);
    for (int j = 0; j < limbs; j++) {
        int64_t bit = 24LL * j;
        uint64_t w = 0;
        size_t i = (size_t)(bit / 32);
        int off = (int)(bit % 32);
        if (i < a.size()) w |= (uint64_t)a[i] >> off;
        if (i + 1 < a.size()) w |= (uint64_t)a[i + 1] << (32 - off);
        out[j] = (double)(w & 0xFFFFFF);
    }
    return out;
}

std::unique_ptr<RnsPlan> build(int count) {
    const Table &t = table();
    auto plan = std::make_unique<RnsPlan>();
    plan->count = count;
    plan->single = std::min(count, t.single);
    plan->planes = plan->single + 2 * (count - plan->single);
    plan->q.assign(t.q.begin(), t.q.begin() + count);
    std::vector<uint32_t> m{1};
    double bits = 0;
    for (uint32_t q : plan->q) {
        mul_small(m, q);
        bits += std::log2((double)q);
    }
    plan->bits = bits;
    // one spare limb beyond M for the signed lift's carries
    plan->limbs = (int)std::ceil((bits + 2) / 24.0) + 1;
    plan->m = limbs24(m, plan->limbs);
    plan->c.resize((size_t)count * plan->limbs);
    plan->inv.resize(count);
    plan->qinv.resize(count);
    // matrix-core CRT table (see RnsPlan)
    int mbits = 0;
    for (size_t w = 0; w < m.size(); w++)
        if (m[w]) mbits = 32 * (int)w + 32 - __builtin_clz(m[w]);
    const int jb = (mbits + 7) / 8;
    plan->crt_jb = jb;
    plan->crt_nt = jb + 2;
    plan->crt_lpad = (count + 15) / 16 * 16;
    plan->crt_npad = (plan->crt_nt + 31) / 32 * 32;
    plan->crt_nu = (8 * jb + 16 + 31) / 32;
    if (4 * plan->crt_nu < plan->crt_nt) plan->crt_nu = (plan->crt_nt + 3) / 4;
    plan->crt_table.assign((size_t)std::max(plan->crt_npad, 0) * 2 * (size_t)std::max(plan->crt_lpad, 0), 0);
    plan->m32.assign((size_t)std::max(plan->crt_nu, 1), 0);
    for (size_t w = 0; w < m.size() && (int)w < plan->crt_nu; w++) plan->m32[w] = m[w];
    {
        const int nu = plan->crt_nu;
        double top = 0;
        for (int u = nu - 1; u >= nu - 3 && u >= 0; u--) top = top * 4294967296.0 + (double)plan->m32[u];
        plan->m_top = top;
    }
    std::vector<uint32_t> mq;
    for (int l = 0; l < count; l++) {
        uint32_t q = plan->q[l];
        div_small(m, q, mq);
        plan->inv[l] = inverse(mod_small(mq, q), q);
        plan->qinv[l] = 1.0 / (double)q;
        auto limbs = limbs24(mq, plan->limbs);
        std::copy(limbs.begin(), limbs.end(), plan->c.begin() + (size_t)l * plan->limbs);
        // balanced base-256 digits of M/q_l (< M: jb bytes, one more for the carry)
        int carry = 0;
        for (int t = 0; t <= jb; t++) {
            const size_t w = (size_t)t / 4;
            int d = (t < jb && w < mq.size() ? (int)((mq[w] >> (8 * (t % 4))) & 255u) : 0) + carry;
            carry = d >= 128 ? 1 : 0;
            if (carry) d -= 256;
            const size_t row = (size_t)2 * plan->crt_lpad;
            plan->crt_table[(size_t)t * row + 2 * l] = (int8_t)d;
            plan->crt_table[(size_t)(t + 1) * row + 2 * l + 1] = (int8_t)d;
        }
    }
    return plan;
}

}  // namespace

const std::vector<uint32_t> &rns_moduli() { return table().q; }
int rns_single_count() { return table().single; }

double rns_needed_bits(int64_t wa, int64_t wb, int64_t k) {
    double lk = k > 1 ? std::ceil(std::log2((double)k)) : 0.0;
    return (double)(wa + wb) + lk + 3.0;
}

const RnsPlan *rns_plan_for_bits(double needed) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<RnsPlan>> cache;
    const Table &t = table();
    double have = 0;
    int count = 0;
    while (have < needed) {
        if (count == (int)t.q.size()) return nullptr;
        have += std::log2((double)t.q[count]);
        count++;
    }
    if (count == 0) count = 1;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(count);
    if (it != cache.end()) return it->second.get();
    auto p = build(count);
    const RnsPlan *raw = p.get();
    cache.emplace(count, std::move(p));
    (void)gcd;
    return raw;
}

}  // namespace sdpx
