// Device-resident multiple-precision LDL^T and sampled arrow factor: the
// same recursion and roundings as ldl.cpp / sampled.cpp (and so as the
// reference, arrow.rs:156-490, 1935-2074; sampled/mod.rs:1836-2143), with
// every panel step as a GPU kernel (mp_kernels.hip) and every exact product
// as an RNS product on device pointers.  The host only walks the recursion
// and launches; values stay in HBM.  A dot whose exponent spread overflows
// the GPU window, or a declined product, sends the whole factorization to the
// host path, which gives the same bits.
#include <cstring>
#include <functional>
#include <vector>

#include "common.h"
#include "ldl.h"
#include "ldl_device.h"
#include "mp_kernels.h"
#include "mpx.h"
#include "rns_gemm.h"
#include "sampled.h"

namespace sdpx {
namespace {

// Kernel-category times of the device pipelines (measurement only): HIP
// event pairs around the launches of each category while profiling is on,
// summed when collected.  0 ldl_dots panels, 1 triangular-solve bases,
// 2 forward rows (Y), 3 exact products (every RNS stage).
struct PhaseLog {
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    double ms[4] = {0, 0, 0, 0};
    int64_t launches[4] = {0, 0, 0, 0};
};
thread_local PhaseLog t_phase;

struct PhaseMark {
    int cat;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    PhaseMark(int c, hipStream_t s) : cat(c), st(s) {
        if (!rns_profiling()) return;
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        HIP_CHECK(hipEventRecord(a, st));
    }
    ~PhaseMark() {
        if (!a) return;
        (void)hipEventRecord(b, st);
        t_phase.marks.push_back({cat, {a, b}});
        t_phase.launches[cat]++;
    }
};

// Workspace slots of the device factor (the ABI's staging uses 72-79).
constexpr int WS_LDL_ARENA = 70, WS_LDL_SCRATCH = 71, WS_LDL_SAVED = 80;

// Bump allocator over one workspace slot; stream order makes reuse after a
// reset safe (all users are on the library stream).
struct Arena {
    uint8_t *base = nullptr;
    size_t size = 0, top = 0;
    void *alloc(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (top + bytes > size) throw Error(SDPX_GPU_DEVICE_ERROR, "device arena exhausted");
        void *p = base + top;
        top += bytes;
        return p;
    }
};

struct DCtx {
    int n;
    size_t vb;
    hipStream_t st;
    Arena *arena;
    const int8_t *d_signs;  // device
    int sign_mode;          // 0 all +1, 1 shared, 2 per pivot
    const void *d_reg_eps, *d_reg_delta;
    int *d_status;  // [0] worst ldl status, [1] regularized, [2] window/decline flag
    uint8_t *ldl_scratch;
    size_t ldl_scratch_bytes;
    LdlStats *stats;
    bool declined = false;
    uint8_t *at(uint8_t *p, size_t idx) const { return p + idx * vb; }
};

// (uploads go through pinned staging, so no wait: arena memory released
// at a reset is only reused by work queued later on the same stream)
int32_t *upload_ints(DCtx &c, const std::vector<size_t> &v) {
    std::vector<int32_t> h(v.begin(), v.end());
    int32_t *d = (int32_t *)c.arena->alloc(sizeof(int32_t) * std::max<size_t>(h.size(), 1));
    upload_async(d, h.data(), sizeof(int32_t) * h.size());
    return d;
}

void product(DCtx &c, char ta, char tb, size_t m, size_t n, size_t k, const uint8_t *a, size_t lda,
             const uint8_t *b, size_t ldb, uint8_t *out, bool packed) {
    if (m == 0 || n == 0) return;
    PhaseMark pm(3, c.st);
    RnsGemmArgs g{c.n, ta, tb, (int64_t)m, (int64_t)n, (int64_t)k, a, (int64_t)lda, b,
                  (int64_t)ldb, packed, packed ? n * (n + 1) / 2 : m * n, out};
    int r = rns_gemm_device(g);
    if (r == SDPX_GPU_DECLINED) {
        c.declined = true;
        return;
    }
    if (r != SDPX_GPU_OK) throw Error(r, "device exact product failed");
    c.stats->gpu_products++;
    c.stats->gpu_mp_macs += (double)k * (packed ? n * (n + 1) / 2.0 : (double)m * n);
}

// Several products with one host wait (rns_gemm_batch); after(i) queues the
// work that consumes item i's output before item i + 1 runs.
void products_batch(DCtx &c, const std::vector<RnsGemmArgs> &items, const std::function<void(size_t)> &after) {
    if (items.empty() || c.declined) return;
    PhaseMark pm(3, c.st);
    size_t failed = 0;
    const int r = rns_gemm_batch(
        items,
        [&](size_t i) {
            const RnsGemmArgs &g = items[i];
            if (g.m > 0 && g.n > 0) {
                c.stats->gpu_products++;
                c.stats->gpu_mp_macs += (double)g.k * (g.upper ? g.n * (g.n + 1) / 2.0 : (double)g.m * g.n);
            }
            if (after) after(i);
        },
        &failed);
    if (r == SDPX_GPU_DECLINED) {
        c.declined = true;
        return;
    }
    if (r != SDPX_GPU_OK) throw Error(r, "device exact product failed");
}

// ldl_dots of one block on the GPU, queued without a wait: a failed pivot
// (status 1) or a dot beyond its window (status 2) stays in d_status[0]
// (the kernels keep the largest code and every later launch returns at
// once), and the caller reads the outcome once when the whole factorization
// is queued.
void dots_device(DCtx &c, uint8_t *l, size_t ld, size_t n, const std::vector<size_t> &first,
                 size_t sign_off, uint8_t *dinv) {
    size_t mark = c.arena->top;
    int32_t *d_first = upload_ints(c, first);
    LdlBlockDesc bd{(int64_t)0, (int64_t)ld, (int64_t)n, 0, 0, (int64_t)sign_off, 0};
    LdlBlockDesc *d_bd = (LdlBlockDesc *)c.arena->alloc(sizeof(LdlBlockDesc));
    upload_async(d_bd, &bd, sizeof bd);
    const size_t need = ldl_scratch_records(n) * c.vb;
    if (need > c.ldl_scratch_bytes) throw Error(SDPX_GPU_DEVICE_ERROR, "ldl scratch too small");
    PhaseMark pm(0, c.st);
    mp_ldl_dots(c.n, l, dinv, d_first, c.d_signs, c.sign_mode, d_bd, 1, (int64_t)n, c.ldl_scratch, c.d_reg_eps,
                c.d_reg_delta, c.d_status, c.st);
    c.arena->top = mark;
}

void solve_t_device(DCtx &c, size_t m, size_t n, uint8_t *L, size_t ldl, uint8_t *B, size_t ldb,
                    const std::vector<size_t> &low) {
    if (m == 0 || n == 0 || c.declined) return;
    if (n <= 128) {
        size_t mark = c.arena->top;
        int32_t *d_low = upload_ints(c, low);
        void *ranges = c.arena->alloc(8 * n);
        void *work = c.arena->alloc(4 * (m + 1));
        PhaseMark pm(1, c.st);
        mp_exp_ranges(c.n, L, ldl, n, 0, ranges, c.st);
        mp_trsm_base(c.n, L, ldl, n, B, ldb, m, d_low, ranges, c.d_status + 2, work, c.st);
        c.arena->top = mark;
        c.stats->base_rows += m;
        return;
    }
    const size_t n1 = n / 2, n2 = n - n / 2;
    std::vector<size_t> low1(m), low2(m);
    for (size_t i = 0; i < m; i++) {
        low1[i] = std::min(low[i], n1);
        low2[i] = low[i] > n1 ? low[i] - n1 : 0;
    }
    solve_t_device(c, m, n1, L, ldl, B, ldb, low1);
    size_t mark = c.arena->top;
    uint8_t *p = (uint8_t *)c.arena->alloc(m * n2 * c.vb);
    product(c, 'N', 'T', m, n2, n1, B, ldb, c.at(L, n1), ldl, p, false);
    if (c.declined) return;
    uint8_t *B2 = c.at(B, n1 * ldb);
    mp_ew_sub_rect(c.n, B2, ldb, p, m, m, n2, c.st);
    c.arena->top = mark;
    solve_t_device(c, m, n2, c.at(L, n1 + n1 * ldl), ldl, B2, ldb, low2);
}

// X L = B (L unit lower n x n, X/B m x n): recursive halves from the right
// with one exact product between them, rows by exact dots below 128 columns.
void solve_n_device(DCtx &c, size_t m, size_t n, uint8_t *L, size_t ldl, uint8_t *B, size_t ldb) {
    if (m == 0 || n == 0 || c.declined) return;
    if (n <= 128) {
        // the ranges buffer is released at the caller's next arena reset,
        // after this launch in stream order
        void *ranges = c.arena->alloc(8 * n);
        PhaseMark pm(1, c.st);
        mp_exp_ranges(c.n, L, ldl, n, 1, ranges, c.st);
        mp_trsm_base_n(c.n, L, ldl, n, B, ldb, m, ranges, c.d_status + 2, c.st);
        return;
    }
    const size_t n1 = n / 2, n2 = n - n / 2;
    uint8_t *B2 = c.at(B, n1 * ldb);
    solve_n_device(c, m, n2, c.at(L, n1 + n1 * ldl), ldl, B2, ldb);
    size_t mark = c.arena->top;
    uint8_t *p = (uint8_t *)c.arena->alloc(m * n1 * c.vb);
    // B1 -= X2 L21  (m x n2 times n2 x n1)
    product(c, 'N', 'N', m, n1, n2, B2, ldb, c.at(L, n1), ldl, p, false);
    if (c.declined) return;
    mp_ew_sub_rect(c.n, B, ldb, p, m, m, n1, c.st);
    c.arena->top = mark;
    solve_n_device(c, m, n1, L, ldl, B, ldb);
}

// ldl_rec on the GPU, queued; the outcome is in d_status (see dots_device)
void rec_device(DCtx &c, uint8_t *l, size_t ld, size_t n, const std::vector<size_t> &first,
                size_t sign_off, uint8_t *dinv) {
    if (c.declined) return;
    if (n <= 128) {
        dots_device(c, l, ld, n, first, sign_off, dinv);
        return;
    }
    size_t low = n, cut = 0;
    bool have = false;
    for (size_t cc = n - 1; cc >= 1; cc--) {
        low = std::min(low, first[cc]);
        auto dist = [&](size_t x) { return x > n / 2 ? x - n / 2 : n / 2 - x; };
        if (low >= cc && (!have || dist(cc) < dist(cut))) {
            cut = cc;
            have = true;
        }
    }
    if (have) {
        std::vector<size_t> f1(first.begin(), first.begin() + cut), tail(n - cut);
        for (size_t i = cut; i < n; i++) tail[i - cut] = first[i] - cut;
        rec_device(c, l, ld, cut, f1, sign_off, dinv);
        rec_device(c, c.at(l, cut + cut * ld), ld, n - cut, tail, sign_off + cut, c.at(dinv, cut));
        return;
    }
    const size_t n1 = n / 2, n2 = n - n / 2;
    std::vector<size_t> f1(first.begin(), first.begin() + n1);
    rec_device(c, l, ld, n1, f1, sign_off, dinv);
    if (c.declined) return;
    size_t mark = c.arena->top;
    uint8_t *l11 = (uint8_t *)c.arena->alloc(n1 * n1 * c.vb);
    HIP_CHECK(hipMemcpy2DAsync(l11, n1 * c.vb, l, ld * c.vb, n1 * c.vb, n1, hipMemcpyDeviceToDevice, c.st));
    std::vector<size_t> lowv(n2);
    for (size_t i = 0; i < n2; i++) lowv[i] = std::min(first[n1 + i], n1);
    uint8_t *a21 = c.at(l, n1);
    solve_t_device(c, n2, n1, l11, n1, a21, ld, lowv);
    if (c.declined) return;
    c.arena->top = mark;  // l11 no longer read (stream order)
    uint8_t *w = (uint8_t *)c.arena->alloc(n2 * n1 * c.vb);
    HIP_CHECK(hipMemcpy2DAsync(w, n2 * c.vb, a21, ld * c.vb, n2 * c.vb, n1, hipMemcpyDeviceToDevice, c.st));
    mp_ew_mul_diag(c.n, a21, ld, dinv, a21, ld, n2, n1, false, c.st);
    uint8_t *packed = (uint8_t *)c.arena->alloc(n2 * (n2 + 1) / 2 * c.vb);
    product(c, 'N', 'T', n2, n2, n1, w, n2, a21, ld, packed, true);
    if (c.declined) return;
    uint8_t *a22 = c.at(l, n1 + n1 * ld);
    mp_ew_sub_packed(c.n, a22, ld, packed, n2, 1, c.st);
    c.arena->top = mark;
    std::vector<size_t> tail(n2);
    for (size_t i = 0; i < n2; i++) tail[i] = first[n1 + i] > n1 ? first[n1 + i] - n1 : 0;
    rec_device(c, a22, ld, n2, tail, sign_off + n1, c.at(dinv, n1));
}

// The envelopes of several square blocks (pointer, ld, order) with one
// read-back.
std::vector<std::vector<size_t>> envelopes_device(DCtx &c, const std::vector<const uint8_t *> &ls,
                                                  const std::vector<size_t> &lds,
                                                  const std::vector<size_t> &ns) {
    size_t mark = c.arena->top, total = 0;
    for (size_t n : ns) total += n;
    int32_t *d = (int32_t *)c.arena->alloc(sizeof(int32_t) * std::max<size_t>(total, 1));
    size_t off = 0;
    for (size_t b = 0; b < ls.size(); b++) {
        mp_envelope(c.n, ls[b], lds[b], ns[b], d + off, c.st);
        off += ns[b];
    }
    std::vector<int32_t> h(total);
    if (total) HIP_CHECK(hipMemcpyAsync(h.data(), d, sizeof(int32_t) * total, hipMemcpyDeviceToHost, c.st));
    HIP_CHECK(hipStreamSynchronize(c.st));
    c.arena->top = mark;
    std::vector<std::vector<size_t>> out;
    off = 0;
    for (size_t n : ns) {
        out.emplace_back(h.begin() + off, h.begin() + off + n);
        off += n;
    }
    return out;
}

std::vector<size_t> envelope_device(DCtx &c, const uint8_t *l, size_t ld, size_t n) {
    size_t mark = c.arena->top;
    int32_t *d = (int32_t *)c.arena->alloc(sizeof(int32_t) * std::max<size_t>(n, 1));
    mp_envelope(c.n, l, ld, n, d, c.st);
    std::vector<int32_t> h(n);
    if (n) HIP_CHECK(hipMemcpyAsync(h.data(), d, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c.st));
    HIP_CHECK(hipStreamSynchronize(c.st));
    c.arena->top = mark;
    return std::vector<size_t>(h.begin(), h.end());
}

struct DeviceSetup {
    Device &dev;
    Arena arena;
    DCtx ctx;
    DeviceSetup(int limbs, size_t arena_bytes, size_t scratch_bytes, const int8_t *signs_host,
                size_t nsigns, const void *reg_eps_host, const void *reg_delta_host, LdlStats *st)
        : dev(device()) {
        const size_t vb = value_bytes(limbs);
        arena.base = (uint8_t *)dev.ws.get(WS_LDL_ARENA, arena_bytes);
        arena.size = arena_bytes;
        arena.top = 0;
        ctx.n = limbs;
        ctx.vb = vb;
        ctx.st = dev.stream;
        ctx.arena = &arena;
        ctx.stats = st;
        ctx.ldl_scratch = (uint8_t *)dev.ws.get(WS_LDL_SCRATCH, scratch_bytes);
        ctx.ldl_scratch_bytes = scratch_bytes;
        // constants: status, signs, regularization values
        ctx.d_status = (int *)arena.alloc(sizeof(int) * 4);
        HIP_CHECK(hipMemsetAsync(ctx.d_status, 0, sizeof(int) * 4, ctx.st));
        ctx.sign_mode = nsigns == 0 ? 0 : (nsigns == 1 ? 1 : 2);
        int8_t *ds = (int8_t *)arena.alloc(std::max<size_t>(nsigns, 1));
        if (nsigns) HIP_CHECK(hipMemcpyAsync(ds, signs_host, nsigns, hipMemcpyHostToDevice, ctx.st));
        ctx.d_signs = ds;
        ctx.d_reg_eps = ctx.d_reg_delta = nullptr;
        if (reg_eps_host) {
            void *e = arena.alloc(vb), *d = arena.alloc(vb);
            HIP_CHECK(hipMemcpyAsync(e, reg_eps_host, vb, hipMemcpyHostToDevice, ctx.st));
            HIP_CHECK(hipMemcpyAsync(d, reg_delta_host, vb, hipMemcpyHostToDevice, ctx.st));
            ctx.d_reg_eps = e;
            ctx.d_reg_delta = d;
        }
        HIP_CHECK(hipStreamSynchronize(ctx.st));
    }
    // {worst ldl code, regularized pivots, window overflow of the row kernels,
    // 1 + a failing block}; code 4 (a bounded wait of the one-launch panel
    // kernel ran out) is a device error
    void status(int (&s)[4]) {
        HIP_CHECK(hipMemcpyAsync(s, ctx.d_status, sizeof s, hipMemcpyDeviceToHost, ctx.st));
        HIP_CHECK(hipStreamSynchronize(ctx.st));
        if (s[0] == 4) throw Error(SDPX_GPU_DEVICE_ERROR, "a panel kernel's bounded wait ran out");
    }
    bool window_overflow() {
        int s[4];
        status(s);
        return s[2] != 0;
    }
};

size_t ldl_arena_bytes(size_t n, size_t vb) {
    // l11 or w copy plus the packed product, nested over halving levels
    return (size_t)(2.5 * (double)n * n * vb) + (64u << 20);
}

// DenseLeaf::factor_loaded (arrow.rs:808-850) on the device, queued and then
// waited for: from order 256 ldl_rec, and when that fails on a pivot (one
// rounding per level can cancel a pivot to an exact zero that exact dots
// keep nonzero, arrow.rs:820-846) the input, saved before the factor, is
// restored and factored by ldl_dots, the failed attempt's regularization
// count dropped; below order 256 ldl_dots.  The envelope comes from the
// input.  Returns 1 ok, 0 failed pivot, -1 declined (window or product).
int factor_loaded_device(DeviceSetup &ds, uint8_t *l, size_t ld, size_t n, uint8_t *dinv,
                         LdlStats &stats, bool recursive) {
    DCtx &c = ds.ctx;
    std::vector<size_t> first = envelope_device(c, l, ld, n);
    int s[4];
    if (recursive) {
        uint8_t *saved = (uint8_t *)ds.dev.ws.get(WS_LDL_SAVED, n * n * c.vb);
        HIP_CHECK(hipMemcpy2DAsync(saved, n * c.vb, l, ld * c.vb, n * c.vb, n, hipMemcpyDeviceToDevice, c.st));
        rec_device(c, l, ld, n, first, 0, dinv);
        ds.status(s);
        // A base block runs only while no window overflow was flagged
        // before it, so a failed pivot (status 1) is the recursive factor's
        // own; what ran after it (products and row kernels on the unfinished
        // factor) is discarded, declines and overflows included.
        if (s[0] != 1) {
            if (c.declined || s[2] != 0 || s[0] == 2) return -1;
            stats.regularized += s[1];
            return 1;
        }
        HIP_CHECK(hipMemcpy2DAsync(l, ld * c.vb, saved, n * c.vb, n * c.vb, n, hipMemcpyDeviceToDevice, c.st));
        HIP_CHECK(hipMemsetAsync(c.d_status, 0, sizeof(int) * 4, c.st));
        c.declined = false;
        stats.retried++;
    }
    const size_t need = ldl_scratch_records(n) * c.vb + 4096;
    if (need > c.ldl_scratch_bytes) {
        c.ldl_scratch = (uint8_t *)ds.dev.ws.get(WS_LDL_SCRATCH, need);
        c.ldl_scratch_bytes = need;
    }
    dots_device(c, l, ld, n, first, 0, dinv);
    ds.status(s);
    if (s[0] == 1) return 0;  // only the dots kernels ran: the failure is the factor's
    if (c.declined || s[2] != 0 || s[0] == 2) return -1;
    stats.regularized += s[1];
    return 1;
}

// The sampled Schur build of every leaf (SampledSchurWorkspace::
// update_with_pool and entry, sampled/mod.rs:1836-2143, 1973-1988):
// V = Rinv Q (every leaf, one host wait), Gram = V^T V (one wait), then the
// entries H_pq = RN(RN(w_p w_q) RN(K_pq^2)) into leaf i's K x K block at
// record loff[i] of d_h, lower triangle, zeros above.
void build_leaves(DCtx &c, const std::vector<SampledLeafDev> &leaves, const std::vector<size_t> &loff, uint8_t *d_h) {
    const size_t vb = c.vb;
    const size_t mark = c.arena->top;
    std::vector<uint8_t *> Vs(leaves.size()), Gs(leaves.size());
    std::vector<RnsGemmArgs> vprod, gram;
    for (size_t i = 0; i < leaves.size(); i++) {
        const auto &lf = leaves[i];
        Vs[i] = (uint8_t *)c.arena->alloc(lf.h * lf.k * vb);
        Gs[i] = (uint8_t *)c.arena->alloc(lf.k * (lf.k + 1) / 2 * vb);
        vprod.push_back(RnsGemmArgs{c.n, 'N', 'N', (int64_t)lf.h, (int64_t)lf.k, (int64_t)lf.h, lf.rinv,
                                    (int64_t)lf.h, lf.q, (int64_t)lf.h, false, (uint64_t)(lf.h * lf.k), Vs[i]});
        gram.push_back(RnsGemmArgs{c.n, 'T', 'N', (int64_t)lf.k, (int64_t)lf.k, (int64_t)lf.h, Vs[i], (int64_t)lf.h,
                                   Vs[i], (int64_t)lf.h, true, (uint64_t)(lf.k * (lf.k + 1) / 2), Gs[i]});
    }
    products_batch(c, vprod, nullptr);
    products_batch(c, gram, [&](size_t i) {
        mp_sampled_entries(c.n, Gs[i], leaves[i].w, (int64_t)leaves[i].k, c.at(d_h, loff[i]), c.st);
    });
    c.arena->top = mark;
}

// arena bytes of build_leaves: V and the packed Gram of every leaf
size_t build_arena_bytes(const std::vector<SampledLeafDev> &leaves, size_t vb) {
    size_t vals = 0;
    for (auto &l : leaves) vals += l.h * l.k + l.k * (l.k + 1) / 2 + 2 * 256 / vb + 2;
    return vals * vb;
}

}  // namespace

void panel_phase_times(double *ms, int64_t *launches, bool reset) {
    for (auto &m : t_phase.marks) {
        HIP_CHECK(hipEventSynchronize(m.second.second));
        float x = 0;
        HIP_CHECK(hipEventElapsedTime(&x, m.second.first, m.second.second));
        t_phase.ms[m.first] += x;
        (void)hipEventDestroy(m.second.first);
        (void)hipEventDestroy(m.second.second);
    }
    t_phase.marks.clear();
    for (int i = 0; i < 4; i++) {
        if (ms) ms[i] = t_phase.ms[i];
        if (launches) launches[i] = t_phase.launches[i];
    }
    if (reset) t_phase = PhaseLog{};
}

int ldl_factor_device(int limbs, uint8_t *d_l, size_t ld, size_t n, const int8_t *signs,
                      size_t nsigns, const void *reg_eps, const void *reg_delta, uint8_t *d_dinv,
                      LdlStats &stats, bool recursive) {
    const size_t vb = value_bytes(limbs);
    const size_t blk = recursive ? std::min<size_t>(n, 128) : n;
    DeviceSetup ds(limbs, ldl_arena_bytes(n, vb), ldl_scratch_records(blk) * vb + 4096, signs, nsigns,
                   reg_eps, reg_delta, &stats);
    return factor_loaded_device(ds, d_l, ld, n, d_dinv, stats, recursive);
}

int ldl_solve_device(int limbs, const uint8_t *d_l, size_t ldl, size_t n, const uint8_t *d_dinv,
                     uint8_t *d_b, size_t ldb, size_t nrhs) {
    const size_t vb = value_bytes(limbs);
    LdlStats ls;
    DeviceSetup ds(limbs, (size_t)(1.5 * (double)nrhs * n * vb) + (size_t)nrhs * n * vb + (64u << 20),
                   4096, nullptr, 0, nullptr, nullptr, &ls);
    DCtx &c = ds.ctx;
    // X = B^T (nrhs x n): L y = b  <=>  y^T L^T = b^T;  L^T x = z  <=>  x^T L = z^T
    uint8_t *X = (uint8_t *)c.arena->alloc(nrhs * n * vb);
    mp_transpose(limbs, d_b, ldb, n, nrhs, X, nrhs, c.st);
    std::vector<size_t> low(nrhs, 0);
    solve_t_device(c, nrhs, n, (uint8_t *)d_l, ldl, X, nrhs, low);
    if (c.declined) return -1;
    mp_ew_mul_diag(limbs, X, nrhs, d_dinv, X, nrhs, nrhs, n, false, c.st);
    solve_n_device(c, nrhs, n, (uint8_t *)d_l, ldl, X, nrhs);
    if (c.declined) return -1;
    mp_transpose(limbs, X, nrhs, nrhs, n, d_b, ldb, c.st);
    HIP_CHECK(hipStreamSynchronize(c.st));
    if (ds.window_overflow()) return -1;
    return 1;
}

int sampled_arrow_solve_device(int limbs, const std::vector<size_t> &ks, size_t t, const uint8_t *d_l_all,
                               const uint8_t *d_dinv_all, const uint8_t *d_y_all, const uint8_t *d_border,
                               const uint8_t *d_border_dinv, size_t nrhs, uint8_t *d_x, size_t ldx) {
    const size_t vb = value_bytes(limbs);
    size_t ksum = 0, kmax = 0;
    std::vector<size_t> loff, koff;
    for (size_t K : ks) {
        loff.push_back(loff.empty() ? 0 : loff.back() + ks[loff.size() - 1] * ks[loff.size() - 1]);
        koff.push_back(ksum);
        ksum += K;
        kmax = std::max(kmax, K);
    }
    if (!mp_sweep_supports((int64_t)kmax) || !mp_sweep_supports((int64_t)t)) return -1;
    if (nrhs == 0) return 1;
    LdlStats ls;
    const size_t pmax = std::max(t, kmax);
    DeviceSetup ds(limbs, (2 * ksum + pmax) * nrhs * vb + 16 * (ks.size() + 1) * sizeof(SweepDescHost) + (64u << 20),
                   4096, nullptr, 0, nullptr, nullptr, &ls);
    DCtx &c = ds.ctx;
    hipStream_t s = c.st;
    uint8_t *Wb = (uint8_t *)c.arena->alloc(std::max<size_t>(ksum, 1) * nrhs * vb);
    uint8_t *Vb = (uint8_t *)c.arena->alloc(std::max<size_t>(ksum, 1) * nrhs * vb);
    uint8_t *P = (uint8_t *)c.arena->alloc(std::max<size_t>(pmax, 1) * nrhs * vb);
    std::vector<SweepDescHost> ld;
    for (size_t l = 0; l < ks.size(); l++)
        ld.push_back(SweepDescHost{(int64_t)loff[l], (int64_t)ks[l], (int64_t)koff[l], (int64_t)ldx, (int64_t)koff[l]});
    SweepDescHost bd{0, (int64_t)t, (int64_t)ksum, (int64_t)ldx, 0};
    SweepDescHost *d_ld = (SweepDescHost *)c.arena->alloc(sizeof(SweepDescHost) * std::max<size_t>(ld.size(), 1));
    SweepDescHost *d_bd = (SweepDescHost *)c.arena->alloc(sizeof(SweepDescHost));
    if (!ld.empty()) upload_async(d_ld, ld.data(), sizeof(SweepDescHost) * ld.size());
    upload_async(d_bd, &bd, sizeof bd);
    // first_many: w = L^-1 b (kept), v = RN(w D^-1), every leaf and column at once
    mp_arrow_forward(limbs, d_l_all, d_dinv_all, d_ld, (int)ld.size(), (int64_t)kmax, (int64_t)nrhs, d_x, Wb, Vb,
                     (int64_t)ksum, s);
    // couple: border rows -= RN(Y_l^T v_l), leaf by leaf
    uint8_t *xb = c.at(d_x, ksum);
    for (size_t l = 0; l < ks.size() && t; l++) {
        product(c, 'T', 'N', t, nrhs, ks[l], c.at((uint8_t *)d_y_all, koff[l] * t), ks[l], c.at(Vb, koff[l]), ksum,
                P, false);
        if (c.declined) return -1;
        mp_ew_sub_rect(limbs, xb, (int64_t)ldx, P, (int64_t)t, (int64_t)t, (int64_t)nrhs, s);
    }
    // trunk: the border factor's sweeps (backward blocks of 32, FORWARD_BLOCK);
    // a border of more than 64 rows spreads each block's updates over the chip
    if (t > 64) {
        mp_sweep_blocked(limbs, d_border, (int64_t)t, (int64_t)nrhs, xb, (int64_t)ldx, false, 32, s);
        mp_ew_mul_diag(limbs, xb, (int64_t)ldx, d_border_dinv, xb, (int64_t)ldx, (int64_t)t, (int64_t)nrhs, true, s);
        mp_sweep_blocked(limbs, d_border, (int64_t)t, (int64_t)nrhs, xb, (int64_t)ldx, true, 32, s);
    } else if (t) {
        mp_arrow_forward(limbs, d_border, d_border_dinv, d_bd, 1, (int64_t)t, (int64_t)nrhs, d_x, nullptr, nullptr, 0,
                         s);
        mp_arrow_backward(limbs, d_border, d_bd, 1, (int64_t)t, (int64_t)nrhs, d_x, 32, s);
    }
    // second_many: x_l = L^-T RN(RN(w - RN(Y_l y)) D^-1) (backward blocks of 16, LEAF_BLOCK)
    for (size_t l = 0; l < ks.size(); l++) {
        const uint8_t *S = P;
        if (t) {
            product(c, 'N', 'N', ks[l], nrhs, t, c.at((uint8_t *)d_y_all, koff[l] * t), ks[l], xb, ldx, P, false);
            if (c.declined) return -1;
        } else {
            mp_zero_values(limbs, P, (int64_t)(ks[l] * nrhs), s);  // no border: Y y is empty
        }
        mp_arrow_second(limbs, c.at(Wb, koff[l]), (int64_t)ksum, S, (int64_t)ks[l], c.at((uint8_t *)d_dinv_all, koff[l]),
                        c.at(d_x, koff[l]), (int64_t)ldx, (int64_t)ks[l], (int64_t)nrhs, s);
    }
    mp_arrow_backward(limbs, d_l_all, d_ld, (int)ld.size(), (int64_t)kmax, (int64_t)nrhs, d_x, 16, s);
    HIP_CHECK(hipStreamSynchronize(s));
    return 1;
}

int sampled_arrow_device(int limbs, const std::vector<SampledLeafDev> &leaves, size_t t,
                         uint8_t *d_l_all, uint8_t *d_dinv_all, uint8_t *d_y_all, uint8_t *d_border,
                         const int8_t *border_signs, size_t nborder_signs, uint8_t *d_border_dinv,
                         SampledStats &st) {
    const size_t vb = value_bytes(limbs);
    size_t kmax = 0, hmax = 0, ksum = 0, build_vals = 0;
    std::vector<size_t> loff, doff, yoff;
    for (auto &l : leaves) {
        kmax = std::max(kmax, l.k);
        hmax = std::max(hmax, l.h);
        build_vals += l.h * l.k + l.k * (l.k + 1) / 2 + 2 * 256 / vb + 2;  // V, Gram (and alignment)
        loff.push_back(loff.empty() ? 0 : loff.back() + leaves[loff.size() - 1].k * leaves[loff.size() - 1].k);
        doff.push_back(ksum);
        yoff.push_back(ksum * t);
        ksum += l.k;
    }
    // leaf scratch, Z of every leaf, one packed contribution, the border factor
    size_t arena = ldl_arena_bytes(t, vb) + (build_vals + ksum * t) * vb +
                   t * (t + 1) / 2 * vb + 16 * ksum + (64u << 20);
    size_t per_leaf = ldl_scratch_records(kmax);
    size_t scratch = std::max<size_t>(leaves.size() * per_leaf * vb, ldl_scratch_records(128) * vb) + 4096;
    LdlStats ls;
    DeviceSetup ds(limbs, arena, scratch, border_signs, nborder_signs, nullptr, nullptr, &ls);
    DCtx &c = ds.ctx;
    hipStream_t s = c.st;
    hipEvent_t ev[4];
    for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(ev[0], s));
    // 1. Schur build of every leaf
    build_leaves(c, leaves, loff, d_l_all);
    HIP_CHECK(hipEventRecord(ev[1], s));
    if (!c.declined) {
        // 2. every leaf factor in one launch, one workgroup per leaf
        std::vector<size_t> first_cat;
        std::vector<LdlBlockDesc> bds;
        std::vector<const uint8_t *> lp;
        std::vector<size_t> lk;
        for (size_t i = 0; i < leaves.size(); i++) {
            lp.push_back(c.at(d_l_all, loff[i]));
            lk.push_back(leaves[i].k);
        }
        const std::vector<std::vector<size_t>> fl = envelopes_device(c, lp, lk, lk);
        for (size_t i = 0; i < leaves.size(); i++) {
            const size_t K = leaves[i].k;
            bds.push_back(LdlBlockDesc{(int64_t)loff[i], (int64_t)K, (int64_t)K, (int64_t)doff[i],
                                       (int64_t)first_cat.size(), 0, (int64_t)(i * per_leaf)});
            first_cat.insert(first_cat.end(), fl[i].begin(), fl[i].end());
        }
        size_t mark = c.arena->top;
        int32_t *d_first = upload_ints(c, first_cat);
        LdlBlockDesc *d_bd = (LdlBlockDesc *)c.arena->alloc(sizeof(LdlBlockDesc) * bds.size());
        upload_async(d_bd, bds.data(), sizeof(LdlBlockDesc) * bds.size());
        int8_t *plus = (int8_t *)c.arena->alloc(1);
        const int8_t one = 1;
        upload_async(plus, &one, 1);
        HIP_CHECK(hipMemsetAsync(c.d_status, 0, sizeof(int) * 2, s));
        {
            PhaseMark pm(0, s);
            mp_ldl_dots(limbs, d_l_all, d_dinv_all, d_first, plus, 1, d_bd, (int)bds.size(), (int64_t)kmax,
                        c.ldl_scratch, nullptr, nullptr, c.d_status, s);
        }
        int status[4];
        ds.status(status);
        c.arena->top = mark;
        if (status[0] == 2) c.declined = true;
        if (status[0] == 1) {
            st.failed_leaf = status[3] - 1;
            for (auto &e : ev) (void)hipEventDestroy(e);
            return 0;
        }
        // 3. Y = L^-1 B of every leaf in one launch and Z = RN(D^-1 Y); then
        // the contributions Y^T Z, each exact and rounded once, subtracted
        // from the border in leaf order
        if (!c.declined) {
            size_t mark3 = c.arena->top;
            int32_t *d_fc = upload_ints(c, first_cat);
            int2 *ranges = (int2 *)c.arena->alloc(sizeof(int2) * std::max<size_t>(ksum, 1));
            std::vector<FwdLeaf> fd;
            for (size_t i = 0; i < leaves.size(); i++) {
                const size_t K = leaves[i].k;
                HIP_CHECK(hipMemcpyAsync(c.at(d_y_all, yoff[i]), leaves[i].b, K * t * vb, hipMemcpyDeviceToDevice,
                                         s));
                mp_exp_ranges(limbs, c.at(d_l_all, loff[i]), (int64_t)K, (int64_t)K, 0, ranges + doff[i], s);
                fd.push_back(FwdLeaf{(int64_t)loff[i], (int64_t)doff[i], (int64_t)K, (int64_t)yoff[i],
                                     (int64_t)doff[i]});
            }
            FwdLeaf *d_fd = (FwdLeaf *)c.arena->alloc(sizeof(FwdLeaf) * fd.size());
            upload_async(d_fd, fd.data(), sizeof(FwdLeaf) * fd.size());
            {
                PhaseMark pm(2, s);
                mp_forward_rows(limbs, d_l_all, d_fc, d_y_all, ranges, d_fd, (int)fd.size(), (int64_t)t,
                                c.d_status + 2, s);
            }
            uint8_t *Z_all = (uint8_t *)c.arena->alloc(std::max<size_t>(ksum, 1) * t * vb);
            for (size_t i = 0; i < leaves.size(); i++)
                mp_ew_mul_diag(limbs, c.at(d_y_all, yoff[i]), leaves[i].k, c.at(d_dinv_all, doff[i]),
                               c.at(Z_all, yoff[i]), leaves[i].k, leaves[i].k, t, true, s);
            // every leaf's product with one host wait; the matrix-core CRT
            // subtracts each rounded entry from the border itself, else the
            // packed product P is subtracted before the next leaf's product
            uint8_t *P = (uint8_t *)c.arena->alloc(t * (t + 1) / 2 * vb);
            std::vector<RnsGemmArgs> contrib;
            std::vector<int> fused(leaves.size(), 0);
            for (size_t i = 0; i < leaves.size(); i++) {
                const int64_t K = (int64_t)leaves[i].k;
                RnsGemmArgs g{limbs, 'T', 'N', (int64_t)t, (int64_t)t, K, c.at(d_y_all, yoff[i]), K,
                              c.at(Z_all, yoff[i]), K, true, (uint64_t)(t * (t + 1) / 2), P};
                g.sub = d_border;
                g.sub_ld = (int64_t)t;
                g.sub_done = &fused[i];
                contrib.push_back(g);
            }
            products_batch(c, contrib, [&](size_t i) {
                if (!fused[i]) mp_ew_sub_packed(limbs, d_border, t, P, t, 0, s);
            });
            c.arena->top = mark3;
        }
    }
    HIP_CHECK(hipEventRecord(ev[2], s));
    int border_result = -1;
    if (!c.declined && !ds.window_overflow()) {
        // factor_border (arrow.rs:3378-3404): the Schur complement mirrored
        // from the upper triangle, factored as factor_loaded does
        mp_mirror_upper(limbs, d_border, t, t, s);
        border_result = factor_loaded_device(ds, d_border, t, t, d_border_dinv, ls, t >= 256);
    }
    HIP_CHECK(hipEventRecord(ev[3], s));
    HIP_CHECK(hipEventSynchronize(ev[3]));
    float a = 0, b = 0, d = 0;
    HIP_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIP_CHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
    HIP_CHECK(hipEventElapsedTime(&d, ev[2], ev[3]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    st.build_s = a / 1e3;
    st.leaves_s = b / 1e3;
    st.border_s = d / 1e3;
    st.gpu_products += ls.gpu_products;
    st.retried += ls.retried;
    return border_result;
}

int sampled_schur_device(int limbs, const std::vector<SampledLeafDev> &leaves, uint8_t *d_h_all) {
    const size_t vb = value_bytes(limbs);
    std::vector<size_t> loff;
    size_t hs = 0;
    for (auto &l : leaves) {
        loff.push_back(hs);
        hs += l.k * l.k;
    }
    LdlStats ls;
    DeviceSetup ds(limbs, build_arena_bytes(leaves, vb) + (64u << 20), 4096, nullptr, 0, nullptr, nullptr, &ls);
    build_leaves(ds.ctx, leaves, loff, d_h_all);
    HIP_CHECK(hipStreamSynchronize(ds.ctx.st));
    return ds.ctx.declined ? -1 : 1;
}

namespace {
enum { WS_RF_A = 160, WS_RF_X, WS_RF_E, WS_RF_P, WS_RF_R, WS_RF_N };
}

int sampled_kkt_solve_device(int limbs, const std::vector<size_t> &ks, size_t t, const uint8_t *d_h_all,
                             const std::vector<const uint8_t *> &d_coupling, const uint8_t *d_c,
                             const uint8_t *d_l_all, const uint8_t *d_dinv_all, const uint8_t *d_y_all,
                             const uint8_t *d_border, const uint8_t *d_border_dinv, const RefineSettings &rs,
                             size_t nrhs, const uint8_t *d_b, size_t ldb, uint8_t *d_x, size_t ldx, int32_t *ok,
                             uint8_t *norms, int64_t *refinements) {
    const size_t vb = value_bytes(limbs);
    size_t ksum = 0, hs = 0;
    std::vector<size_t> koff, hoff;
    for (size_t K : ks) {
        koff.push_back(ksum);
        hoff.push_back(hs);
        ksum += K;
        hs += K * K;
    }
    const size_t n = ksum + t, kk = n + nrhs;
    if (refinements) *refinements = 0;
    if (nrhs == 0 || n == 0) return 1;
    Device &dev = device();
    hipStream_t s = dev.stream;
    auto at = [vb](const uint8_t *p, size_t i) { return (uint8_t *)p + i * vb; };
    // [K | B]: the unshifted arrow KKT (n x n, both triangles) with the
    // right-hand sides as nrhs more columns, so that with X = [-p; e_c] one
    // exact product gives every row's b_ic - sum_j K_ij p_j rounded once
    uint8_t *A = (uint8_t *)dev.ws.get(WS_RF_A, n * kk * vb);
    uint8_t *X = (uint8_t *)dev.ws.get(WS_RF_X, kk * nrhs * vb);
    uint8_t *E = (uint8_t *)dev.ws.get(WS_RF_E, n * nrhs * vb);  // residual of column c at c n
    uint8_t *P = (uint8_t *)dev.ws.get(WS_RF_P, n * nrhs * vb);  // corrections, then trial points
    uint8_t *R = (uint8_t *)dev.ws.get(WS_RF_R, n * nrhs * vb);  // product output
    uint8_t *dn = (uint8_t *)dev.ws.get(WS_RF_N, (nrhs + 1) * vb);
    uint8_t *one_dev = at(dn, nrhs);
    std::vector<uint8_t> one(vb, 0);
    ((uint64_t *)one.data())[limbs - 1] = 1ULL << 63;
    *(int32_t *)(one.data() + 8 * limbs) = mpx::KIND_REG;
    *(int32_t *)(one.data() + 8 * limbs + 4) = 1;
    upload_async(one_dev, one.data(), vb);
    mp_zero_values(limbs, A, (int64_t)(n * kk), s);
    for (size_t l = 0; l < ks.size(); l++) {
        const size_t K = ks[l], o = koff[l];
        if (K == 0) continue;
        HIP_CHECK(hipMemcpy2DAsync(at(A, o + o * n), n * vb, at(d_h_all, hoff[l]), K * vb, K * vb, K,
                                   hipMemcpyDeviceToDevice, s));
        mp_mirror_lower(limbs, at(A, o + o * n), (int64_t)K, (int64_t)n, s);
        if (t) {
            HIP_CHECK(hipMemcpy2DAsync(at(A, o + ksum * n), n * vb, d_coupling[l], K * vb, K * vb, t,
                                       hipMemcpyDeviceToDevice, s));
            mp_transpose(limbs, d_coupling[l], (int64_t)K, (int64_t)K, (int64_t)t, at(A, ksum + o * n), (int64_t)n, s);
        }
    }
    if (t) {
        HIP_CHECK(hipMemcpy2DAsync(at(A, ksum + ksum * n), n * vb, d_c, t * vb, t * vb, t, hipMemcpyDeviceToDevice, s));
        mp_mirror_upper(limbs, at(A, ksum + ksum * n), (int64_t)t, (int64_t)n, s);
    }
    HIP_CHECK(hipMemcpy2DAsync(at(A, n * n), n * vb, d_b, ldb * vb, n * vb, nrhs, hipMemcpyDeviceToDevice, s));

    std::vector<uint8_t> hn(nrhs * vb), tmp(nrhs * vb);
    auto fetch_norms = [&](size_t m) {
        HIP_CHECK(hipMemcpyAsync(tmp.data(), dn, m * vb, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    };
    // ExactRows::residual of points Q[:, j] (ld ldq) for columns cols[j]:
    // E[:, cols[j]] = RN(b_c - K Q[:, j]), their infinity norms into hn;
    // false when the product declines
    auto residual = [&](const uint8_t *Q, size_t ldq, const std::vector<size_t> &cols) -> bool {
        const size_t m = cols.size();
        mp_zero_values(limbs, X, (int64_t)(kk * m), s);
        for (size_t j = 0; j < m; j++) {
            mp_axpby(limbs, at(Q, j * ldq), (int64_t)ldq, (int64_t)n, 1, 2, nullptr, 0, nullptr, at(X, j * kk),
                     (int64_t)kk, 0, false, false, s);
            HIP_CHECK(hipMemcpyAsync(at(X, n + cols[j] + j * kk), one_dev, vb, hipMemcpyDeviceToDevice, s));
        }
        RnsGemmArgs g{limbs, 'N', 'N', (int64_t)n, (int64_t)m, (int64_t)kk, A, (int64_t)n, X, (int64_t)kk,
                      false, (uint64_t)(n * m), R};
        const int r = rns_gemm_device(g);
        if (r == SDPX_GPU_DECLINED) return false;
        if (r != SDPX_GPU_OK) throw Error(r, "residual product failed");
        for (size_t j = 0; j < m; j++)
            HIP_CHECK(hipMemcpyAsync(at(E, cols[j] * n), at(R, j * n), n * vb, hipMemcpyDeviceToDevice, s));
        mp_col_absmax(limbs, R, (int64_t)n, (int64_t)n, (int64_t)m, dn, s);
        fetch_norms(m);
        for (size_t j = 0; j < m; j++) std::memcpy(&hn[cols[j] * vb], &tmp[j * vb], vb);
        return true;
    };
    auto view = [&](const std::vector<uint8_t> &v, size_t c) { return mpx::View{limbs, v.data() + c * vb}; };

    // x = K^-1 b (DirectLDLKKTSolver::solve_many), then the refinement
    HIP_CHECK(hipMemcpy2DAsync(d_x, ldx * vb, d_b, ldb * vb, n * vb, nrhs, hipMemcpyDeviceToDevice, s));
    if (sampled_arrow_solve_device(limbs, ks, t, d_l_all, d_dinv_all, d_y_all, d_border, d_border_dinv, nrhs, d_x,
                                   ldx) < 0)
        return -1;
    mp_col_absmax(limbs, d_b, (int64_t)ldb, (int64_t)n, (int64_t)nrhs, dn, s);
    fetch_norms(nrhs);
    std::vector<uint8_t> tol(nrhs * vb), last(nrhs * vb), xnorm(nrhs * vb), t1(vb), ratio(vb);
    for (size_t c = 0; c < nrhs; c++) {  // abstol + reltol * normb
        mpx::mul(mpx::View{limbs, rs.reltol}, mpx::View{limbs, &tmp[c * vb]}, t1.data());
        mpx::add(mpx::View{limbs, rs.abstol}, mpx::View{limbs, t1.data()}, &tol[c * vb]);
    }
    std::vector<size_t> all(nrhs);
    for (size_t c = 0; c < nrhs; c++) all[c] = c;
    if (!residual(d_x, ldx, all)) return -1;
    xnorm = hn;
    std::vector<char> okv(nrhs), active(nrhs);
    for (size_t c = 0; c < nrhs; c++) okv[c] = active[c] = view(hn, c).finite();
    const mpx::View stop{limbs, rs.stop_ratio}, unit{limbs, one.data()};
    int64_t nref = 0;
    for (uint32_t it = 0; it < rs.max_iter; it++) {
        std::vector<size_t> act;
        for (size_t c = 0; c < nrhs; c++) {
            if (active[c] && mpx::cmp(view(hn, c), view(tol, c)) <= 0) active[c] = 0;
            if (active[c]) act.push_back(c);
        }
        if (act.empty()) break;
        last = hn;
        nref += (int64_t)act.size();
        const size_t m = act.size();
        for (size_t j = 0; j < m; j++)
            HIP_CHECK(hipMemcpyAsync(at(P, j * n), at(E, act[j] * n), n * vb, hipMemcpyDeviceToDevice, s));
        if (sampled_arrow_solve_device(limbs, ks, t, d_l_all, d_dinv_all, d_y_all, d_border, d_border_dinv, m, P,
                                       n) < 0)
            return -1;
        for (size_t j = 0; j < m; j++)  // the prospective solution x + dx
            mp_axpby(limbs, at(d_x, act[j] * ldx), (int64_t)ldx, (int64_t)n, 1, 1, nullptr, 1, nullptr, at(P, j * n),
                     (int64_t)n, 0, false, false, s);
        if (!residual(P, n, act)) return -1;
        for (size_t j = 0; j < m; j++) {
            const size_t c = act[j];
            if (!view(hn, c).finite()) {
                okv[c] = active[c] = 0;
                continue;
            }
            mpx::div(view(last, c), view(hn, c), ratio.data());
            const mpx::View r{limbs, ratio.data()};
            if (mpx::cmp(r, stop) < 0) {
                active[c] = 0;
                if (mpx::cmp(r, unit) <= 0) continue;
            }
            HIP_CHECK(hipMemcpyAsync(at(d_x, c * ldx), at(P, j * n), n * vb, hipMemcpyDeviceToDevice, s));
            std::memcpy(&xnorm[c * vb], &hn[c * vb], vb);
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    for (size_t c = 0; c < nrhs; c++) {
        if (ok) ok[c] = okv[c];
        if (norms) std::memcpy(norms + c * vb, &xnorm[c * vb], vb);
    }
    if (refinements) *refinements = nref;
    return 1;
}

}  // namespace sdpx
