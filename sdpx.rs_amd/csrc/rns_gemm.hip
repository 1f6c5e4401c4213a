// Exact residue-number-system (RNS) products of multiple-precision matrices
// on gfx950 INT8 matrix cores.
//
// Contract (crates/solver/src/algebra/dense/blas/rns_blas.rs:1-21,
// 2134-2417): out = op(A) op(B) with every entry the exact dot product
// rounded once, nearest-even, at 64N bits.  The result does not depend on
// moduli, windows or tiling, so this pipeline is free to choose its own:
//
//   1. scan    per row of op(A) / column of op(B): the lowest set bit
//              position (exponent - 64N + trailing zeros), the exponent range
//              and non-finite flags.  Each row/column gets its own binary
//              window, so a' = A_ik 2^-lo_i is an integer of at most
//              w_a = max_i (emax_i - lo_i) bits.
//   2. encode  a' -> symmetric residues mod q_l, written as signed-byte
//              digit planes [plane][row][k] (k contiguous, zero padded).
//   3. product per modulus and 128x128 output tile, INT8 MFMA
//              (v_mfma_i32_32x32x32_i8) with exact INT32 accumulation,
//              then x_l = (S mod q_l) (M/q_l)^-1 mod q_l as u16.
//   4. crt     S = sum_l x_l (M/q_l) - round(sum_l x_l/q_l) M in 24-bit
//              limbs (exact FP64 arithmetic), then one nearest-even rounding
//              to 64N bits at scale 2^(lo_i + lo_j).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "common.h"
#include "mpdev.h"
#include "rns_gemm.h"
#include "rns_plan.h"

namespace sdpx {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int TILE = 128;        // output tile edge of the product kernel
constexpr int KSTEP = 64;        // inner-dimension bytes per LDS stage
constexpr int LDS_ROW = KSTEP;   // LDS row (bytes) of one staged operand row
// 16-byte chunk c of staged row r sits at chunk c ^ ((r >> 2) & 3): the 16
// lanes of a b128 access (4 rows x 4 chunks on the store side, 16 rows x 1
// chunk on the MFMA fragment side) then hit 16 distinct 16-byte banks.
__device__ __forceinline__ int lds_off(int r, int byte) {
    return r * LDS_ROW + (((byte >> 4) ^ ((r >> 2) & 3)) << 4);
}
constexpr int KCHUNK = 32768;    // inner elements per exact INT32 accumulation
constexpr int XPAD = 130;        // padded LDS column of the transposed x tile

// ---------------------------------------------------------------------------
// 1. scan
// ---------------------------------------------------------------------------
struct ScanAcc {
    int lo, emin, emax, flags;  // flags: 1 non-finite, 2 has a non-zero
};

__device__ __forceinline__ void scan_value(const uint32_t *rec, int n, ScanAcc &s) {
    int kind = (int)rec[2 * n];
    int ak = kind < 0 ? -kind : kind;
    if (ak == KIND_ZERO) return;
    if (ak != KIND_REGULAR) {
        s.flags |= 1;
        return;
    }
    int e = (int)rec[2 * n + 1];
    int tz = 0;
    for (int w = 0; w < 2 * n; w++) {
        uint32_t x = rec[w];
        if (x) {
            tz += __builtin_ctz(x);
            break;
        }
        tz += 32;
    }
    s.lo = min(s.lo, e - 64 * n + tz);
    s.emin = min(s.emin, e);
    s.emax = max(s.emax, e);
    s.flags |= 2;
}

__device__ __forceinline__ void scan_merge(ScanAcc &a, const ScanAcc &b) {
    a.lo = min(a.lo, b.lo);
    a.emin = min(a.emin, b.emin);
    a.emax = max(a.emax, b.emax);
    a.flags |= b.flags;
}

// Vectors contiguous in memory (sv == 1): a block takes 64 vectors x 4
// position lanes over one chunk of positions (blockIdx.y: chunks of pch);
// the chunks merge into out by atomic min / max / or (out set to the empty
// accumulator first by scan_init), so the grid spans the positions of a few
// long vectors as well as many short ones.
__global__ void scan_init(int64_t V, int4 *__restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = make_int4(INT_MAX, INT_MAX, INT_MIN, 0);
}
__global__ __launch_bounds__(256) void scan_vec_fast(const uint32_t *__restrict__ data, int n,
                                                     int64_t V, int64_t K, int64_t sv, int64_t sp,
                                                     int64_t pch, int4 *__restrict__ out) {
    __shared__ ScanAcc red[4][64];
    int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    int64_t v = (int64_t)blockIdx.x * 64 + tx;
    const int64_t p0 = (int64_t)blockIdx.y * pch, p1 = min(K, p0 + pch);
    ScanAcc s{INT_MAX, INT_MAX, INT_MIN, 0};
    size_t W = 2 * (size_t)n + 2;
    if (v < V)
        for (int64_t p = p0 + ty; p < p1; p += 4) scan_value(data + (v * sv + p * sp) * W, n, s);
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && v < V) {
        for (int t = 1; t < 4; t++) scan_merge(s, red[t][tx]);
        int *o = (int *)(out + v);
        if (s.lo != INT_MAX) atomicMin(o, s.lo);
        if (s.emin != INT_MAX) atomicMin(o + 1, s.emin);
        if (s.emax != INT_MIN) atomicMax(o + 2, s.emax);
        if (s.flags) atomicOr(o + 3, s.flags);
    }
}

// Positions contiguous in memory (sp == 1): one wave per vector.
__global__ __launch_bounds__(256) void scan_pos_fast(const uint32_t *__restrict__ data, int n,
                                                     int64_t V, int64_t K, int64_t sv, int64_t sp,
                                                     int4 *__restrict__ out) {
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t v = (int64_t)blockIdx.x * 4 + w;
    if (v >= V) return;
    ScanAcc s{INT_MAX, INT_MAX, INT_MIN, 0};
    size_t W = 2 * (size_t)n + 2;
    for (int64_t p = lane; p < K; p += 64) scan_value(data + (v * sv + p * sp) * W, n, s);
    for (int off = 32; off > 0; off >>= 1) {
        ScanAcc o{__shfl_xor(s.lo, off), __shfl_xor(s.emin, off), __shfl_xor(s.emax, off),
                  __shfl_xor(s.flags, off)};
        scan_merge(s, o);
    }
    if (lane == 0) out[v] = make_int4(s.lo, s.emin, s.emax, s.flags);
}

// ---------------------------------------------------------------------------
// 2. encode
// ---------------------------------------------------------------------------
// Bits [b, b + 24) of the mantissa (2n words, LSB first); outside reads 0.
__device__ __forceinline__ uint32_t bits24(const uint32_t *m, int words, int b) {
    if (b <= -24 || b >= 32 * words) return 0;
    if (b < 0) return (m[0] << (-b)) & 0xFFFFFFu;
    int w = b >> 5, off = b & 31;
    uint64_t x = m[w];
    if (w + 1 < words) x |= (uint64_t)m[w + 1] << 32;
    return (uint32_t)(x >> off) & 0xFFFFFFu;
}

// Records of a 32 x 32 (vector x position) tile are staged in LDS with
// coalesced runs along whichever index is contiguous in memory; each thread
// then encodes entries of one position column so the digit-plane writes are
// contiguous along k.
template <int TJMAX>
__global__ __launch_bounds__(256) void encode_kernel(
    const uint32_t *__restrict__ data, int n, int64_t V, int64_t K, int64_t sv, int64_t sp,
    const int4 *__restrict__ scan, int TJ, const double *__restrict__ pow2,
    const uint32_t *__restrict__ mq, const double *__restrict__ mqinv, int L, int single,
    int8_t *__restrict__ planes, int64_t plane_stride, int64_t k_pad) {
    extern __shared__ uint32_t lds[];
    const int W = 2 * n + 2;
    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.y * 32, p0 = (int64_t)blockIdx.x * 32;
    const bool runs_along_p = (sp == 1);
    for (int r = 0; r < 32; r++) {
        int64_t run_len;
        const uint32_t *src;
        if (runs_along_p) {
            int64_t v = v0 + r;
            run_len = v < V ? min((int64_t)32, max((int64_t)0, K - p0)) : 0;
            src = data + (size_t)(v * sv + p0) * W;
        } else {
            int64_t p = p0 + r;
            run_len = p < K ? min((int64_t)32, max((int64_t)0, V - v0)) : 0;
            src = data + (size_t)(v0 + p * sp) * W;
        }
        for (int w = tid; w < 32 * W; w += 256) {
            int idx = w / W, word = w - idx * W;
            uint32_t val = (idx < run_len) ? src[w] : (word == 2 * n ? (uint32_t)KIND_ZERO : 0u);
            int slot = runs_along_p ? (r * 32 + idx) : (idx * 32 + r);
            lds[slot * W + word] = val;
        }
    }
    __syncthreads();
    const int pp = tid & 31;
    for (int u = 0; u < 4; u++) {
        const int vv = (tid >> 5) + 8 * u;
        const int64_t v = v0 + vv, p = p0 + pp;
        const uint32_t *rec = lds + (vv * 32 + pp) * W;
        const int kind = (int)rec[2 * n];
        const bool nz = (kind == KIND_REGULAR || kind == -KIND_REGULAR);
        const bool negative = kind < 0;
        int s = 0;
        if (nz) s = (int)rec[2 * n + 1] - 64 * n - scan[v].x;
        uint32_t ch[TJMAX];
#pragma unroll
        for (int j = 0; j < TJMAX; j++) ch[j] = (nz && j < TJ) ? bits24(rec, 2 * n, 24 * j - s) : 0u;
        const int64_t dst = v * k_pad + p;
        for (int l = 0; l < L; l++) {
            double acc = 0.0;
            const double *pw = pow2 + (size_t)l * TJ;
#pragma unroll
            for (int j = 0; j < TJMAX; j++)
                if (j < TJ) acc = fma((double)ch[j], pw[j], acc);
            const int q = (int)mq[l];
            const double qd = (double)q;
            double quo = floor(acc * mqinv[l]);
            double r = fma(-quo, qd, acc);
            if (r < 0) r += qd;
            if (r >= qd) r -= qd;
            int ri = (int)r;
            if (negative && ri) ri = q - ri;
            if (ri >= ((q + 1) >> 1)) ri -= q;
            if (l < single) {
                planes[(size_t)l * plane_stride + dst] = (int8_t)ri;
            } else {
                int g = single + 2 * (l - single);
                int lo8 = ((ri + 128) & 255) - 128;
                int hi8 = (ri - lo8) >> 8;
                planes[(size_t)g * plane_stride + dst] = (int8_t)hi8;
                planes[(size_t)(g + 1) * plane_stride + dst] = (int8_t)lo8;
            }
        }
    }
}

// Encode on the INT8 matrix cores.  With X = sum_b x_b 256^b the aligned
// magnitude of a value (B_pad bytes) and T_bl = 256^b mod q_l taken
// symmetric and split into signed bytes T = 256 th + tl,
//   S_l = sum_b (x_b - 128) (256 th_bl + tl_bl)  ==  X - C_l  (mod q_l),
// C_l = 128 sum_b T_bl.  Per 32 values x 32 moduli that is two 32x32x32
// MFMAs per 32 bytes; E_l in [2^30, 2^30 + q_l) with E_l == C_l makes
// 256 GH + GL + E_l a positive 31-bit integer, reduced by a Barrett step
// (floor(2^32 / q) estimate, one correction).  A workgroup encodes tiles of
// 4 vectors x 64 positions (wave w: vector v0 + w, two 32-position M-tiles)
// one block of 32 moduli at a time, staging the digit bytes in LDS so the
// plane writes are 16-byte runs along k.  With few tiles the blocks of a
// tile run side by side instead (blockIdx.y: the block, gridDim.y of them,
// each workgroup holding only its block's table rows).
constexpr int ENC_TV = 4, ENC_TP = 64;
constexpr int ENC_ROW = ENC_TV * ENC_TP + 16;  // LDS bytes per staged plane

__global__ __launch_bounds__(256) void encode_mfma_kernel(
    const uint32_t *__restrict__ data, int n, int64_t V, int64_t K, int64_t sv, int64_t sp,
    const int4 *__restrict__ scan, int nk, const int8_t *__restrict__ th, const int8_t *__restrict__ tl,
    const uint32_t *__restrict__ econst, const uint32_t *__restrict__ mq, const uint32_t *__restrict__ mbar,
    int L, int single, int8_t *__restrict__ planes, int64_t plane_stride, int64_t k_pad,
    int64_t tiles_v, int64_t tiles_p) {
    extern __shared__ __attribute__((aligned(16))) int8_t esm[];
    const int bpad = 32 * nk, trow = bpad + 16;  // table row: padded for b128 reads
    const int nblocks = (L + 31) / 32;
    const int split = gridDim.y > 1;                        // one block per workgroup
    const int srows = split ? 32 : 32 * nblocks, row0 = split ? 32 * (int)blockIdx.y : 0;
    int8_t *sth = esm, *stl = esm + srows * trow;
    int8_t *sout = esm + 2 * srows * trow;  // [64 staged planes][ENC_ROW]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, col = lane & 31;
    const int W32 = 2 * n + 2;
    // table rows -> LDS (zero rows for padded moduli)
    for (int e = tid; e < srows * (bpad / 4); e += 256) {
        const int lr = e / (bpad / 4), w4 = e - lr * (bpad / 4), l = row0 + lr;
        uint32_t vh = 0, vl = 0;
        if (l < L) {
            vh = ((const uint32_t *)th)[(size_t)l * (bpad / 4) + w4];
            vl = ((const uint32_t *)tl)[(size_t)l * (bpad / 4) + w4];
        }
        *(uint32_t *)(sth + lr * trow + 4 * w4) = vh;
        *(uint32_t *)(stl + lr * trow + 4 * w4) = vl;
    }
    __syncthreads();
    const int64_t ntiles = tiles_v * tiles_p;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t v0 = (tile / tiles_p) * ENC_TV, p0 = (tile % tiles_p) * ENC_TP;
        const int64_t v = v0 + wave;
        // A fragments (value bytes ^ 0x80) and sign masks of the two M-tiles
        v4i af[2][3];
        uint32_t negmask[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int64_t p = p0 + 32 * t + col;
            const uint32_t *rec = nullptr;
            int kind = KIND_ZERO, s = 0;
            if (v < V && p < K) {
                rec = data + (size_t)(v * sv + p * sp) * W32;
                kind = (int)rec[2 * n];
                if (kind == KIND_REGULAR || kind == -KIND_REGULAR)
                    s = (int)rec[2 * n + 1] - 64 * n - scan[v].x;
                else
                    rec = nullptr;
            }
            negmask[t] = (uint32_t)__ballot(rec != nullptr && kind < 0);
            // X = M 2^s: dword w of X is bits [32 w - s, 32 w - s + 32) of M
            const int sw = s >= 0 ? s / 32 : -((-s + 31) / 32);  // floor(s / 32)
            const int sb = s - 32 * sw;                            // [0, 32)
            const int off = (32 - sb) & 31, adj = sb > 0 ? 1 : 0;
            auto word = [&](int idx) -> uint32_t {
                return (rec && idx >= 0 && idx < 2 * n) ? rec[idx] : 0u;
            };
#pragma unroll
            for (int ks = 0; ks < 3; ks++) {
                if (ks < nk) {
                    const int w0 = 8 * ks + 4 * h;
                    const int i0 = w0 - sw - adj;
                    uint32_t m[5];
#pragma unroll
                    for (int u = 0; u < 5; u++) m[u] = word(i0 + u);
                    uint32_t d[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) d[u] = __builtin_amdgcn_alignbit(m[u + 1], m[u], off) ^ 0x80808080u;
                    af[t][ks] = v4i{(int)d[0], (int)d[1], (int)d[2], (int)d[3]};
                } else {
                    af[t][ks] = v4i{0, 0, 0, 0};
                }
            }
        }
        for (int nb = blockIdx.y; nb < nblocks; nb += gridDim.y) {
            const int c = 32 * nb + col;  // this lane's modulus
            const int pfirst = 32 * nb <= single ? 32 * nb : single + 2 * (32 * nb - single);
            const int clast = min(L, 32 * nb + 32);
            const int pend = clast <= single ? clast : single + 2 * (clast - single);
            uint32_t q = 1, mqc = 0, ec = 0;
            if (c < L) {
                q = mq[c];
                mqc = mbar[c];
                ec = econst[c];
            }
            const int8_t *bh = sth + (c - row0) * trow + 16 * h, *bl = stl + (c - row0) * trow + 16 * h;
            __syncthreads();  // staged planes of the previous block are out
#pragma unroll
            for (int t = 0; t < 2; t++) {
                v16i gh = v16i{0}, gl = v16i{0};
#pragma unroll
                for (int ks = 0; ks < 3; ks++) {
                    if (ks < nk) {
                        const v4i fh = *(const v4i *)(bh + 32 * ks);
                        const v4i fl = *(const v4i *)(bl + 32 * ks);
                        gh = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[t][ks], fh, gh, 0, 0, 0);
                        gl = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[t][ks], fl, gl, 0, 0, 0);
                    }
                }
                if (c < L) {
                    uint32_t lo4[4] = {0, 0, 0, 0}, hi4[4] = {0, 0, 0, 0};
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int row = 8 * (r >> 2) + 4 * h + (r & 3);
                        const uint32_t x = (uint32_t)(256 * gh[r] + gl[r]) + ec;
                        uint32_t rr = x - __umulhi(x, mqc) * q;
                        if (rr >= q) rr -= q;
                        if (((negmask[t] >> row) & 1u) && rr) rr = q - rr;
                        int ri = (int)rr;
                        if (ri >= (int)((q + 1) >> 1)) ri -= (int)q;
                        const int sh = 8 * (r & 3);
                        if (c < single) {
                            lo4[r >> 2] |= (uint32_t)(ri & 255) << sh;
                        } else {
                            const int l8 = ((ri + 128) & 255) - 128;
                            hi4[r >> 2] |= (uint32_t)(((ri - l8) >> 8) & 255) << sh;
                            lo4[r >> 2] |= (uint32_t)(l8 & 255) << sh;
                        }
                    }
                    const int lp = (c < single ? c : single + 2 * (c - single)) - pfirst;
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        const int pos = 32 * t + 8 * g + 4 * h;
                        if (c < single) {
                            *(uint32_t *)(sout + lp * ENC_ROW + wave * ENC_TP + pos) = lo4[g];
                        } else {
                            *(uint32_t *)(sout + lp * ENC_ROW + wave * ENC_TP + pos) = hi4[g];
                            *(uint32_t *)(sout + (lp + 1) * ENC_ROW + wave * ENC_TP + pos) = lo4[g];
                        }
                    }
                }
            }
            __syncthreads();
            // staged planes [pfirst, pend) -> HBM, 16-byte runs along k
            const int chunks = (pend - pfirst) * ENC_TV * (ENC_TP / 16);
            for (int e = tid; e < chunks; e += 256) {
                const int lp = e / (ENC_TV * 4), rest = e - lp * ENC_TV * 4;
                const int tv = rest >> 2, part = rest & 3;
                const v4i val = *(const v4i *)(sout + lp * ENC_ROW + tv * ENC_TP + 16 * part);
                *(v4i *)(planes + (size_t)(pfirst + lp) * plane_stride + (size_t)(v0 + tv) * k_pad +
                         p0 + 16 * part) = val;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 3. product
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mod_q(double v, double qd, double qinv) {
    double quo = floor(v * qinv);
    double r = fma(-quo, qd, v);
    if (r < 0) r += qd;
    if (r >= qd) r -= qd;
    return (int)r;
}

// x_l in [0, q) is stored as its symmetric representative x' (x' == x mod
// q, -q/2 < x' <= q/2) in balanced base-256 digits, lo = low byte, hi = high
// byte, x' = 256 hi + lo with both digits in [-128, 127]: the CRT's matrix
// product consumes the two bytes directly as INT8 operands.
__device__ __forceinline__ uint16_t x_encode(int x, int q) {
    const int xs = x >= ((q + 1) >> 1) ? x - q : x;
    const int lo = ((xs + 128) & 255) - 128;
    const int hi = (xs - lo) >> 8;
    return (uint16_t)((lo & 255) | ((hi & 255) << 8));
}
// back to [0, q)
__device__ __forceinline__ int x_decode(uint16_t u, int q) {
    const int x = (int)(int8_t)(u & 255) + 256 * (int)(int8_t)(u >> 8);
    return x < 0 ? x + q : x;
}

// One modulus x one 128x128 output tile per workgroup; 4 waves in 2x2, each
// 64x64 = 2x2 MFMA blocks of 32x32.  D = digits per residue (1 or 2).
// Planes of modulus l start at plane index plane_of(l); for D = 2 the high
// digit plane comes first.
template <int D>
__global__ __launch_bounds__(256) void product_kernel(
    const int8_t *__restrict__ A, const int8_t *__restrict__ B, int64_t a_plane, int64_t b_plane,
    int64_t k_pad, int64_t k_begin, int64_t k_end, const uint32_t *__restrict__ mq,
    const uint32_t *__restrict__ minv, const double *__restrict__ mqinv, int single,
    int mod_first, int mod_count, const int2 *__restrict__ tiles, int ntiles, int tiles_m,
    int tiles_n, int accumulate, uint16_t *__restrict__ X) {
    constexpr int PLANE_TILE = TILE * LDS_ROW;  // bytes of one plane tile in LDS
    constexpr int MAIN_BYTES = 2 * D * PLANE_TILE;
    constexpr int XT_BYTES = TILE * XPAD * 2;
    constexpr int LDS_BYTES = MAIN_BYTES > XT_BYTES ? MAIN_BYTES : XT_BYTES;
    __shared__ __attribute__((aligned(16))) int8_t smem[LDS_BYTES];
    int8_t *sA = smem;                      // [D][128][LDS_ROW]
    int8_t *sB = smem + D * PLANE_TILE;     // [D][128][LDS_ROW]

    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs,
    // so slot b runs on XCD b % 8.  The grid is padded to a multiple of 8 and
    // each XCD takes a contiguous range of (modulus, tile) work, so tiles
    // sharing operand panels meet in one L2.
    const int total = mod_count * ntiles;
    const int per = (int)(gridDim.x >> 3);
    const int bid = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (bid >= total) return;
    const int l = mod_first + bid / ntiles;
    const int2 tile = tiles[bid % ntiles];
    const int tm = tile.x, tn = tile.y;
    const int plane = l < single ? l : single + 2 * (l - single);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    const int8_t *gA[D], *gB[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        gA[d] = A + (size_t)(plane + d) * a_plane + (size_t)tm * TILE * k_pad;
        gB[d] = B + (size_t)(plane + d) * b_plane + (size_t)tn * TILE * k_pad;
    }
    // global -> register staging: 2 x 16 B per thread per plane tile
    const int c0 = tid, c1 = tid + 256;
    const int r0 = c0 >> 2, o0 = (c0 & 3) * 16, r1 = c1 >> 2, o1 = (c1 & 3) * 16;
    v4i ra[D][2], rb[D][2];
    auto load = [&](int64_t k) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            ra[d][0] = *(const v4i *)(gA[d] + (size_t)r0 * k_pad + k + o0);
            ra[d][1] = *(const v4i *)(gA[d] + (size_t)r1 * k_pad + k + o1);
            rb[d][0] = *(const v4i *)(gB[d] + (size_t)r0 * k_pad + k + o0);
            rb[d][1] = *(const v4i *)(gB[d] + (size_t)r1 * k_pad + k + o1);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int d = 0; d < D; d++) {
            *(v4i *)(sA + d * PLANE_TILE + lds_off(r0, o0)) = ra[d][0];
            *(v4i *)(sA + d * PLANE_TILE + lds_off(r1, o1)) = ra[d][1];
            *(v4i *)(sB + d * PLANE_TILE + lds_off(r0, o0)) = rb[d][0];
            *(v4i *)(sB + d * PLANE_TILE + lds_off(r1, o1)) = rb[d][1];
        }
    };

    constexpr int NACC = D == 1 ? 1 : 3;  // D=2: hh, hl+lh, ll
    v16i acc[NACC][2][2];
#pragma unroll
    for (int a = 0; a < NACC; a++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[a][i][j] = v16i{0};

    const int frag_row = lane & 31, frag_k = (lane >> 5) * 16;
    load(k_begin);
    for (int64_t k = k_begin; k < k_end; k += KSTEP) {
        __syncthreads();
        store();
        __syncthreads();
        if (k + KSTEP < k_end) load(k + KSTEP);
#pragma unroll
        for (int kk = 0; kk < KSTEP; kk += 32) {
            v4i fa[D][2], fb[D][2];
#pragma unroll
            for (int d = 0; d < D; d++)
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    fa[d][i] = *(const v4i *)(sA + d * PLANE_TILE +
                                              lds_off(wm * 64 + i * 32 + frag_row, kk + frag_k));
                    fb[d][i] = *(const v4i *)(sB + d * PLANE_TILE +
                                              lds_off(wn * 64 + i * 32 + frag_row, kk + frag_k));
                }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if constexpr (D == 1) {
                        acc[0][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[0][i], fb[0][j],
                                                                             acc[0][i][j], 0, 0, 0);
                    } else {
                        acc[0][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[0][i], fb[0][j],
                                                                             acc[0][i][j], 0, 0, 0);
                        acc[1][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[0][i], fb[1][j],
                                                                             acc[1][i][j], 0, 0, 0);
                        acc[1][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[1][i], fb[0][j],
                                                                             acc[1][i][j], 0, 0, 0);
                        acc[2][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[1][i], fb[1][j],
                                                                             acc[2][i][j], 0, 0, 0);
                    }
                }
        }
    }

    // residues -> x_l = S (M/q)^-1 mod q, transposed through LDS
    const int q = (int)mq[l];
    const double qd = (double)q, qinv = mqinv[l];
    const double inv = (double)minv[l];
    const double r16 = (double)((65536 % q)), r8 = (double)(256 % q);
    // x = S (M/q)^-1 mod q in one reduction: the digit weights 2^16, 2^8 and
    // (M/q)^-1 folded into w16, w8, w1 < q; with |acc| <= 2^30 and q < 2^16
    // the weighted sum stays below 2^47, exact in binary64
    const double w1 = inv, w8 = (double)mod_q(r8 * inv, qd, qinv), w16 = (double)mod_q(r16 * inv, qd, qinv);
    __syncthreads();
    uint16_t *sX = (uint16_t *)smem;  // [128 cols][XPAD] (row index fastest)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int x = D == 1 ? mod_q((double)acc[0][i][j][r] * w1, qd, qinv)
                                     : mod_q(fma((double)acc[0][i][j][r], w16,
                                                 fma((double)acc[D == 1 ? 0 : 1][i][j][r], w8,
                                                     (double)acc[D == 1 ? 0 : 2][i][j][r] * w1)),
                                             qd, qinv);
                int row = wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                int col = wn * 64 + j * 32 + (lane & 31);
                sX[col * XPAD + row] = x_encode(x, q);
            }
    __syncthreads();
    uint16_t *gX = X + ((size_t)(l * tiles_n + tn) * tiles_m + tm) * (TILE * TILE);
#pragma unroll
    for (int it = 0; it < 8; it++) {
        int e = (it * 256 + tid) * 8;
        int col = e >> 7, row = e & 127;
        const uint32_t *src = (const uint32_t *)(sX + col * XPAD + row);
        uint32_t w0 = src[0], w1 = src[1], w2 = src[2], w3 = src[3];
        if (accumulate) {
            v4i old = *(const v4i *)(gX + e);
            uint32_t o[4] = {(uint32_t)old.x, (uint32_t)old.y, (uint32_t)old.z, (uint32_t)old.w};
            uint32_t nw[4] = {w0, w1, w2, w3};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int lo = x_decode((uint16_t)(nw[t] & 0xFFFF), q) + x_decode((uint16_t)(o[t] & 0xFFFF), q);
                const int hi = x_decode((uint16_t)(nw[t] >> 16), q) + x_decode((uint16_t)(o[t] >> 16), q);
                nw[t] = (uint32_t)x_encode(lo >= q ? lo - q : lo, q) |
                        ((uint32_t)x_encode(hi >= q ? hi - q : hi, q) << 16);
            }
            w0 = nw[0];
            w1 = nw[1];
            w2 = nw[2];
            w3 = nw[3];
        }
        *(v4i *)(gX + e) = v4i{(int)w0, (int)w1, (int)w2, (int)w3};
    }
}

// Second product kernel: BM x 128 output tiles (BM = 256 for one-digit
// moduli: every wave owns a (BM/2) x 64 block of MI x 2 MFMA tiles, so each
// LDS fragment feeds MI or 2 MFMAs), double-buffered LDS with one barrier per
// 64-byte k stage, and the next stage's global loads in flight during the
// current stage's MFMAs.  Output x planes keep the 128 x 128 tile layout
// (BM / 128 tiles per workgroup).  tiles2 lists (row block of BM, column tile).
template <int D, int BM>
__global__ __launch_bounds__(256) void product2_kernel(
    const int8_t *__restrict__ A, const int8_t *__restrict__ B, int64_t a_plane, int64_t b_plane,
    int64_t k_pad, int64_t k_begin, int64_t k_end, const uint32_t *__restrict__ mq,
    const uint32_t *__restrict__ minv, const double *__restrict__ mqinv, int single,
    int mod_first, int mod_count, const int2 *__restrict__ tiles2, int ntiles2, int tiles_m,
    int tiles_n, int accumulate, uint16_t *__restrict__ X) {
    constexpr int MI = BM / 64;                          // 32-row MFMA blocks per wave
    constexpr int ROWS = BM + TILE;                      // staged rows per digit
    constexpr int STAGE = D * ROWS * LDS_ROW;            // bytes per LDS stage
    constexpr int CH = D * ROWS * (KSTEP / 16) / 256;    // 16-byte chunks per thread
    constexpr int XROW = BM + 2;                         // u16 per staged x column
    extern __shared__ __attribute__((aligned(16))) int8_t psm[];

    const int total = mod_count * ntiles2;
    const int per = (int)(gridDim.x >> 3);
    const int bid = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (bid >= total) return;
    const int l = mod_first + bid / ntiles2;
    const int2 tile = tiles2[bid % ntiles2];
    const int tm2 = tile.x, tn = tile.y;
    const int plane = l < single ? l : single + 2 * (l - single);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    // chunk c: digit d = c / (ROWS * 4), row r = (c / 4) % ROWS, 16-byte part c % 4
    const int8_t *src[CH];
    int dst[CH];
#pragma unroll
    for (int u = 0; u < CH; u++) {
        const int c = tid + 256 * u;
        const int d = c / (ROWS * 4), rr = (c / 4) % ROWS, part = c & 3;
        const int8_t *base = rr < BM ? A + (size_t)(plane + d) * a_plane + ((size_t)tm2 * BM + rr) * k_pad
                                     : B + (size_t)(plane + d) * b_plane + ((size_t)tn * TILE + rr - BM) * k_pad;
        src[u] = base + 16 * part;
        dst[u] = d * ROWS * LDS_ROW + lds_off(rr, 16 * part);
    }
    v4i stg[CH];
    auto gload = [&](int64_t k) {
#pragma unroll
        for (int u = 0; u < CH; u++) stg[u] = *(const v4i *)(src[u] + k);
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < CH; u++) *(v4i *)(psm + buf * STAGE + dst[u]) = stg[u];
    };

    constexpr int NACC = D == 1 ? 1 : 3;
    v16i acc[NACC][MI][2];
#pragma unroll
    for (int a = 0; a < NACC; a++)
#pragma unroll
        for (int i = 0; i < MI; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[a][i][j] = v16i{0};

    const int frag_row = lane & 31, frag_k = (lane >> 5) * 16;
    gload(k_begin);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (int64_t k = k_begin; k < k_end; k += KSTEP) {
        const bool more = k + KSTEP < k_end;
        if (more) gload(k + KSTEP);
        const int8_t *sb = psm + buf * STAGE;
#pragma unroll
        for (int kk = 0; kk < KSTEP; kk += 32) {
            v4i fa[D][MI], fb[D][2];
#pragma unroll
            for (int d = 0; d < D; d++) {
#pragma unroll
                for (int i = 0; i < MI; i++)
                    fa[d][i] = *(const v4i *)(sb + d * ROWS * LDS_ROW +
                                              lds_off(wm * (BM / 2) + i * 32 + frag_row, kk + frag_k));
#pragma unroll
                for (int j = 0; j < 2; j++)
                    fb[d][j] = *(const v4i *)(sb + d * ROWS * LDS_ROW +
                                              lds_off(BM + wn * 64 + j * 32 + frag_row, kk + frag_k));
            }
#pragma unroll
            for (int i = 0; i < MI; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if constexpr (D == 1) {
                        acc[0][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[0][i], fb[0][j], acc[0][i][j], 0, 0, 0);
                    } else {
                        acc[0][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[0][i], fb[0][j], acc[0][i][j], 0, 0, 0);
                        acc[1][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[0][i], fb[1][j], acc[1][i][j], 0, 0, 0);
                        acc[1][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[1][i], fb[0][j], acc[1][i][j], 0, 0, 0);
                        acc[2][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[1][i], fb[1][j], acc[2][i][j], 0, 0, 0);
                    }
                }
        }
        if (more) lstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    // residues -> x_l = S (M/q)^-1 mod q (balanced digits), transposed through LDS
    const int q = (int)mq[l];
    const double qd = (double)q, qinv = mqinv[l];
    const double inv = (double)minv[l];
    const double r16 = (double)((65536 % q)), r8 = (double)(256 % q);
    // x = S (M/q)^-1 mod q in one reduction: the digit weights 2^16, 2^8 and
    // (M/q)^-1 folded into w16, w8, w1 < q; with |acc| <= 2^30 and q < 2^16
    // the weighted sum stays below 2^47, exact in binary64
    const double w1 = inv, w8 = (double)mod_q(r8 * inv, qd, qinv), w16 = (double)mod_q(r16 * inv, qd, qinv);
    uint16_t *sX = (uint16_t *)psm;  // [128 cols][XROW] (row index fastest)
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int x = D == 1 ? mod_q((double)acc[0][i][j][r] * w1, qd, qinv)
                                     : mod_q(fma((double)acc[0][i][j][r], w16,
                                                 fma((double)acc[D == 1 ? 0 : 1][i][j][r], w8,
                                                     (double)acc[D == 1 ? 0 : 2][i][j][r] * w1)),
                                             qd, qinv);
                const int row = wm * (BM / 2) + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = wn * 64 + j * 32 + (lane & 31);
                sX[col * XROW + row] = x_encode(x, q);
            }
    __syncthreads();
#pragma unroll
    for (int sub = 0; sub < BM / TILE; sub++) {
        const int tm = tm2 * (BM / TILE) + sub;
        uint16_t *gX = X + ((size_t)(l * tiles_n + tn) * tiles_m + tm) * (TILE * TILE);
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int e = (it * 256 + tid) * 8;
            const int col = e >> 7, row = e & 127;
            const uint32_t *srcx = (const uint32_t *)(sX + col * XROW + sub * TILE + row);
            uint32_t nw[4] = {srcx[0], srcx[1], srcx[2], srcx[3]};
            if (accumulate) {
                const v4i old = *(const v4i *)(gX + e);
                const uint32_t o[4] = {(uint32_t)old.x, (uint32_t)old.y, (uint32_t)old.z, (uint32_t)old.w};
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int lo = x_decode((uint16_t)(nw[t] & 0xFFFF), q) + x_decode((uint16_t)(o[t] & 0xFFFF), q);
                    const int hi = x_decode((uint16_t)(nw[t] >> 16), q) + x_decode((uint16_t)(o[t] >> 16), q);
                    nw[t] = (uint32_t)x_encode(lo >= q ? lo - q : lo, q) |
                            ((uint32_t)x_encode(hi >= q ? hi - q : hi, q) << 16);
                }
            }
            *(v4i *)(gX + e) = v4i{(int)nw[0], (int)nw[1], (int)nw[2], (int)nw[3]};
        }
    }
}

// ---------------------------------------------------------------------------
// 4. crt + rounding
// ---------------------------------------------------------------------------
template <int JMAX, int BS>
__global__ __launch_bounds__(BS) void crt_kernel(
    const uint16_t *__restrict__ X, int L, const int2 *__restrict__ tiles, int tiles_m,
    int tiles_n, const double *__restrict__ C, const double *__restrict__ Mt,
    const double *__restrict__ qinv, const uint32_t *__restrict__ mq, int J,
    const int4 *__restrict__ scan_a,
    const int4 *__restrict__ scan_b, int64_t m, int64_t nn, int upper, int packed, int n,
    uint32_t *__restrict__ out) {
    __shared__ uint32_t limbs_s[JMAX * BS];
    const int tid = threadIdx.x;
    const int2 tile = tiles[blockIdx.y];
    const int e = blockIdx.x * BS + tid;
    const int il = e & 127, jl = e >> 7;
    const int64_t i = (int64_t)tile.x * TILE + il, j = (int64_t)tile.y * TILE + jl;
    if (i >= m || j >= nn || (upper && i > j)) return;

    const uint16_t *xp = X + ((size_t)tile.y * tiles_m + tile.x) * (TILE * TILE) + e;
    const size_t xstride = (size_t)tiles_n * tiles_m * (TILE * TILE);
    double acc[JMAX];
#pragma unroll
    for (int t = 0; t < JMAX; t++) acc[t] = 0.0;
    double f = 0.0;
    for (int l = 0; l < L; l++) {
        const double x = (double)x_decode(xp[(size_t)l * xstride], (int)mq[l]);
        f = fma(x, qinv[l], f);
        const double *cl = C + (size_t)l * J;
#pragma unroll
        for (int t = 0; t < JMAX; t++)
            if (t < J) acc[t] = fma(x, cl[t], acc[t]);
    }
    const double tq = rint(f);
    const double two24 = 16777216.0, inv24 = 1.0 / 16777216.0;
    double carry = 0.0;
    bool any = false;
#pragma unroll
    for (int t = 0; t < JMAX; t++) {
        if (t < J) {
            double v = fma(-tq, Mt[t], acc[t]) + carry;
            double hi = floor(v * inv24);
            double lo = fma(-hi, two24, v);
            carry = hi;
            uint32_t li = (uint32_t)lo;
            limbs_s[t * BS + tid] = li;
            any |= li != 0;
        }
    }
    const bool negative = carry < 0;
    uint32_t *out_rec;
    {
        int64_t o = packed ? (j * (j + 1) / 2 + i) : (i + j * m);
        out_rec = out + (size_t)o * (2 * n + 2);
    }
    if (!any) {
        for (int w = 0; w < 2 * n; w++) out_rec[w] = 0;
        out_rec[2 * n] = KIND_ZERO;
        out_rec[2 * n + 1] = 0;
        return;
    }
    if (negative) {  // two's complement negation over J limbs
        uint32_t c = 1;
        for (int t = 0; t < J; t++) {
            uint32_t v = ((~limbs_s[t * BS + tid]) & 0xFFFFFFu) + c;
            c = v >> 24;
            limbs_s[t * BS + tid] = v & 0xFFFFFFu;
        }
    }
    int top = J - 1;
    while (top > 0 && limbs_s[top * BS + tid] == 0) top--;
    const int bits = 24 * top + (32 - __builtin_clz(limbs_s[top * BS + tid]));
    const int P = 64 * n;
    int shift = bits - P;
    auto bit_word = [&](int b) -> uint32_t {  // bits [b, b + 32)
        uint32_t w = 0;
        int t0 = b >= 0 ? b / 24 : -((-b + 23) / 24);
        for (int t = t0; t <= t0 + 2; t++) {
            if (t < 0 || t >= J) continue;
            uint32_t v = limbs_s[t * BS + tid];
            int pos = 24 * t - b;
            if (pos >= 32 || pos <= -24) continue;
            w |= pos >= 0 ? (v << pos) : (v >> (-pos));
        }
        return w;
    };
    uint32_t mant[2 * MAX_LIMBS];
    for (int u = 0; u < 2 * n; u++) mant[u] = bit_word(shift + 32 * u);
    if (shift > 0) {
        const int rb = shift - 1;
        const bool round = (limbs_s[(rb / 24) * BS + tid] >> (rb % 24)) & 1;
        bool sticky = false;
        for (int t = 0; t < rb / 24 && !sticky; t++) sticky = limbs_s[t * BS + tid] != 0;
        if (!sticky && (rb % 24)) sticky = (limbs_s[(rb / 24) * BS + tid] & ((1u << (rb % 24)) - 1)) != 0;
        if (round && (sticky || (mant[0] & 1))) {
            uint32_t c = 1;
            for (int u = 0; u < 2 * n && c; u++) {
                mant[u] += 1;
                c = mant[u] == 0;
            }
            if (c) {
                mant[2 * n - 1] = 0x80000000u;
                shift += 1;
            }
        }
    }
    const int64_t scale = (int64_t)scan_a[i].x + scan_b[j].x;
    const int64_t expo = scale + shift + P;
    if (expo > (int64_t)((1 << 30) - 1)) {
        for (int w = 0; w < 2 * n; w++) out_rec[w] = 0;
        out_rec[2 * n] = (uint32_t)(negative ? -KIND_INF : KIND_INF);
        out_rec[2 * n + 1] = 0;
        return;
    }
    if (expo < (int64_t)(1 - (1 << 30))) {
        for (int w = 0; w < 2 * n; w++) out_rec[w] = 0;
        out_rec[2 * n] = KIND_ZERO;
        out_rec[2 * n + 1] = 0;
        return;
    }
    for (int u = 0; u < 2 * n; u++) out_rec[u] = mant[u];
    out_rec[2 * n] = (uint32_t)(negative ? -KIND_REGULAR : KIND_REGULAR);
    out_rec[2 * n + 1] = (uint32_t)(int32_t)expo;
}

// CRT on the INT8 matrix cores.  The product kernel leaves x'_l as balanced
// digit pairs (lo, hi), so V = sum_l x'_l (M/q_l) = sum_t 256^t R_t with
//   R_t = sum_l lo_l c_{l,t} + hi_l c_{l,t-1}
// (c the balanced base-256 digits of M/q_l): one INT8 product of the
// (nt x 2L) table with the (2L x outputs) digit matrix, exact in INT32.  A
// workgroup (2 waves) takes one 128-row column of an output tile and stages
// its x planes in LDS as they lie ([l][output], rows padded for the
// fragment reads).  In the MFMA result a lane holds, for one output, the
// four consecutive byte positions of every other 32-bit limb; one
// cross-half shuffle per limb gives each lane all limbs of one output.  That
// lane then carries, removes k M with k = round(V / M) from the top limbs,
// and rounds the signed result once, nearest-even, to 64N bits.
constexpr int CRT_MAXW = 16;        // mantissa words handled (up to 512 bits)
constexpr size_t CRT_NU_MAX = 40;   // words of M the matrix-core CRT reads (zero-padded)

// The fused subtraction of the arrow contributions: srec = RN(srec - c) for
// the rounded entry c (words mant, kind, exponent) in registers.
template <int NL>
__device__ __forceinline__ void crt_sub_store(uint32_t *srec, const uint32_t *mant, int32_t kind, int32_t expo) {
    constexpr int W = 2 * NL;
    mpd::Val<W> c, sv, r;
#pragma unroll
    for (int w = 0; w < W; w++) c.m[w] = mant[w];
    c.kind = kind;
    c.e = expo;
    mpd::load(srec, sv);
    mpd::add_signed(sv, c, true, r);
    mpd::store(srec, r);
}

template <int NUMAX, int NBMAX, int NL>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(NUMAX <= 24 ? 3 : 2))) void crt_mfma_kernel(
    const uint16_t *__restrict__ X, int L, int lpad, const int2 *__restrict__ tiles, int ntiles,
    int tiles_m, int tiles_n, const int8_t *__restrict__ table, int nt, int npad,
    const uint32_t *__restrict__ mlimbs, int nu, double m_top, const int4 *__restrict__ scan_a,
    const int4 *__restrict__ scan_b, int64_t m, int64_t nn, int upper, int packed, int n_arg,
    uint32_t *__restrict__ out, int chunks, int64_t chunk_stride, uint32_t *__restrict__ sub, int64_t sub_ld) {
    const int n = NL ? NL : n_arg;  // output limbs (a constant in the common widths)
    extern __shared__ __attribute__((aligned(16))) int8_t csm[];
    const int trow = 2 * lpad + 16;  // bytes per table row
    int8_t *stab = csm;
    uint32_t *sl = (uint32_t *)(csm + (size_t)npad * trow);  // [NUMAX][128]: each thread its column
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane >> 5, col = lane & 31;
    const int nks = lpad / 16, nbl = npad / 32;
    const int c16 = 2 * lpad / 16;
    for (int e = tid; e < npad * c16; e += 128) {
        const int t = e / c16, c = e - t * c16;
        *(v4i *)(stab + t * trow + 16 * c) = *(const v4i *)(table + (size_t)t * 2 * lpad + 16 * c);
    }
    __syncthreads();  // the table is in (the only barrier: everything else is per wave or per thread)
    const size_t xstride = (size_t)tiles_n * tiles_m * (TILE * TILE);
    const int64_t items = (int64_t)ntiles * TILE;
    const int P = 64 * n;
    for (int64_t item = blockIdx.x; item < items; item += gridDim.x) {
        const int2 tile = tiles[item / TILE];
        const int jl = (int)(item % TILE);
        const int64_t j = (int64_t)tile.y * TILE + jl, i0 = (int64_t)tile.x * TILE;
        if (j >= nn || (upper && i0 > j)) continue;  // uniform over the workgroup
        const uint16_t *xb = X + ((size_t)tile.y * tiles_m + tile.x) * (TILE * TILE) + (size_t)jl * TILE;
        // B fragments of both output tiles (mt = 0, 1) for every k-slice,
        // read once from the planes; every table block reuses them.  Planes
        // L .. lpad - 1 are allocated but never written: their table
        // coefficients are zero, so whatever they hold adds nothing.  The
        // item's address is wave-uniform: one buffer resource per k-slice (16
        // planes above it) in scalar registers, this lane's byte offset (its
        // half's 8 planes, its output) in a vector register and each plane's
        // offset in a scalar one, so no load computes an address on the
        // vector unit.
        constexpr int NKS = 2 * NBMAX;
        v4i fb[2][NKS];
        const uint64_t xbu = (uint64_t)(uintptr_t)xb;
        const uint64_t xbs = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(xbu >> 32)) << 32) |
                             (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)xbu);
        const uint32_t pstride = (uint32_t)(xstride * sizeof(uint16_t));  // bytes between planes
        const uint32_t loff = 8 * (uint32_t)h * pstride + (uint32_t)((64 * wave + col) * sizeof(uint16_t));
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            fb[0][ks] = fb[1][ks] = v4i{0, 0, 0, 0};
            if (ks < nks) {
                const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(uintptr_t)(xbs + (uint64_t)(16 * ks) * pstride), (short)0, (int)0xFFFFFFFF, 0x00020000);
#pragma unroll
                for (int mt = 0; mt < 2; mt++) {
                    const int vo = (int)(loff + 32 * mt * (uint32_t)sizeof(uint16_t));
                    uint32_t d[4];
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        d[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rs, vo, (int)(2 * u * pstride), 0) |
                               ((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rs, vo, (int)((2 * u + 1) * pstride), 0)
                                << 16);
                    fb[mt][ks] = v4i{(int)d[0], (int)d[1], (int)d[2], (int)d[3]};
                }
            }
        }
        // partial limb sum_b R_{32 nb + 8 g + 4 h + b} 256^b is limb 8 nb + 2 g + h
        // of output (mt, o); lane h owns output o = 64 wave + 32 h + col, i.e.
        // tile mt = h.  Each table block multiplies both tiles; the two halves
        // then trade the partials of the tile the other half owns, and every
        // lane carries its own output's limbs, one 32-row block at a time.
        uint32_t lim[NUMAX];
#pragma unroll
        for (int u = 0; u < NUMAX; u++) lim[u] = 0;
        int64_t carry = 0;
#pragma unroll
        for (int nb = 0; nb < NBMAX; nb++) {
            v16i acc0 = v16i{0}, acc1 = v16i{0};  // blocks past nbl add zeros and carry on
#pragma unroll
            for (int ks = 0; ks < NKS; ks++) {
                if (nb < nbl && ks < nks) {
                    const v4i fa = *(const v4i *)(stab + (32 * nb + col) * trow + 32 * ks + 16 * h);
                    acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa, fb[0][ks], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa, fb[1][ks], acc1, 0, 0, 0);
                }
            }
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int64_t pv0 = (int64_t)acc0[4 * g] + (int64_t)acc0[4 * g + 1] * 256 +
                                    (int64_t)acc0[4 * g + 2] * 65536 + (int64_t)acc0[4 * g + 3] * 16777216;
                const int64_t pv1 = (int64_t)acc1[4 * g] + (int64_t)acc1[4 * g + 1] * 256 +
                                    (int64_t)acc1[4 * g + 2] * 65536 + (int64_t)acc1[4 * g + 3] * 16777216;
                const int64_t got = __shfl_xor(h ? pv0 : pv1, 32, 64);  // the other half's partial of my tile
                const int64_t own = h ? pv1 : pv0;
                const int64_t p0 = h ? got : own, p1 = h ? own : got;  // limbs 2g, 2g + 1
                const int u = 8 * nb + 2 * g;
                if (u < NUMAX) {
                    const int64_t a = carry + p0;
                    lim[u] = (uint32_t)a;
                    carry = a >> 32;
                }
                if (u + 1 < NUMAX) {
                    const int64_t a = carry + p1;
                    lim[u + 1] = (uint32_t)a;
                    carry = a >> 32;
                }
            }
        }
#pragma unroll
        for (int u = 8 * NBMAX; u < NUMAX; u++) {  // the carry above the partial limbs
            const int64_t a = carry;
            lim[u] = (uint32_t)a;
            carry = a >> 32;
        }
        const int64_t i = i0 + 64 * wave + lane;
        const bool active = i < m && !(upper && i > j);
        // V's words in this thread's LDS column: the three top limbs below
        // word nu are read back at the run-time index
#pragma unroll
        for (int u = 0; u < NUMAX; u++) sl[u * 128 + tid] = lim[u];
        const uint32_t l2 = sl[(nu - 1) * 128 + tid], l1 = nu >= 2 ? sl[(nu - 2) * 128 + tid] : 0u,
                       l0 = nu >= 3 ? sl[(nu - 3) * 128 + tid] : 0u;
        // k = round(V / M); S = V - k M over all NUMAX words (M's words
        // zero-padded to NUMAX; S is small, so the words above nu carry its
        // sign)
        // (the same top min(nu, 3) words as m_top, so plans of one or two
        // words scale alike)
        double vt = (double)(int32_t)l2;
        if (nu >= 2) vt = vt * 4294967296.0 + (double)l1;
        if (nu >= 3) vt = vt * 4294967296.0 + (double)l0;
        const int32_t kq = (int32_t)rint(vt / m_top);
        int64_t br = 0;
#pragma unroll
        for (int u = 0; u < NUMAX; u++) {
            const int64_t t = (int64_t)lim[u] - (int64_t)kq * (int64_t)mlimbs[u] + br;
            lim[u] = (uint32_t)t;
            br = t >> 32;
        }
        const bool negative = (int32_t)lim[NUMAX - 1] < 0;
        if (negative) {
            uint64_t c = 1;
#pragma unroll
            for (int u = 0; u < NUMAX; u++) {
                const uint64_t t = (uint64_t)(~lim[u]) + c;
                lim[u] = (uint32_t)t;
                c = t >> 32;
            }
        }
        int top = -1, low = NUMAX;
        uint32_t tw = 0;
#pragma unroll
        for (int u = 0; u < NUMAX; u++) {
            if (lim[u] != 0) {
                top = u;
                tw = lim[u];
            }
            const int d = NUMAX - 1 - u;
            if (lim[d] != 0) low = d;
        }
        // the magnitude's words over V's (each thread its own column)
#pragma unroll
        for (int u = 0; u < NUMAX; u++) sl[u * 128 + tid] = lim[u];
        if (!active) continue;
        uint32_t *rec = out + (size_t)(packed ? j * (j + 1) / 2 + i : i + j * m) * (2 * n + 2);
        uint32_t *srec = nullptr;  // fused subtraction: the target entry
        if constexpr (NL > 0) {
            if (sub) srec = sub + (size_t)(i + j * sub_ld) * (2 * NL + 2);
        }
        if (top < 0 && srec) continue;  // RN(s - 0) = s
        if (top < 0) {
            for (int cidx = 0; cidx < (chunks > 0 ? chunks : 1); cidx++) {
                uint32_t *crec = chunks > 0 ? out + ((size_t)cidx * chunk_stride + (size_t)(i + j * m)) * (2 * n + 2)
                                            : rec;
                for (int w = 0; w < 2 * n; w++) crec[w] = 0;
                crec[2 * n] = KIND_ZERO;
                crec[2 * n + 1] = 0;
            }
            continue;
        }
        auto word = [&](int u) -> uint32_t { return u >= 0 && u < NUMAX ? sl[u * 128 + tid] : 0u; };
        const int bits = 32 * top + 32 - __builtin_clz(tw);
        if (chunks > 0) {
            // expansion form: chunk c holds bits [bits - (c + 1) P, bits - c P)
            // of the magnitude, exactly (no rounding), each normalized on its
            // own; the chunks sum to the exact value
            for (int cidx = 0; cidx < chunks; cidx++) {
                uint32_t *crec = out + ((size_t)cidx * chunk_stride + (size_t)(i + j * m)) * (2 * n + 2);
                const int sc = bits - (cidx + 1) * P;
                const int q = sc >> 5, rs = sc & 31;
                uint32_t mc[CRT_MAXW];
                int hw = -1;
#pragma unroll
                for (int w = 0; w < CRT_MAXW; w++) {
                    mc[w] = w < 2 * n ? __builtin_amdgcn_alignbit(word(q + w + 1), word(q + w), rs) : 0u;
                    if (w < 2 * n && mc[w] != 0) hw = w;
                }
                if (hw < 0) {
                    for (int w = 0; w < 2 * n; w++) crec[w] = 0;
                    crec[2 * n] = KIND_ZERO;
                    crec[2 * n + 1] = 0;
                    continue;
                }
                const int lz = 32 * (2 * n - 1 - hw) + __builtin_clz(mc[hw]);
                const int ws = lz >> 5, bs = lz & 31;
                const int64_t expo = (int64_t)scan_a[i].x + scan_b[j].x + sc - lz + P;
                if (expo > (int64_t)((1 << 30) - 1) || expo < (int64_t)(1 - (1 << 30))) {
                    const bool inf = expo > 0;
                    for (int w = 0; w < 2 * n; w++) crec[w] = 0;
                    crec[2 * n] = (uint32_t)(inf ? (negative ? -KIND_INF : KIND_INF) : KIND_ZERO);
                    crec[2 * n + 1] = 0;
                    continue;
                }
                for (int w = 2 * n - 1; w >= 0; w--) {
                    const int src = w - ws;
                    const uint32_t hi = src >= 0 ? mc[src] : 0u, lo = src >= 1 ? mc[src - 1] : 0u;
                    crec[w] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
                }
                crec[2 * n] = (uint32_t)(negative ? -KIND_REGULAR : KIND_REGULAR);
                crec[2 * n + 1] = (uint32_t)(int32_t)expo;
            }
            continue;
        }
        int shift = bits - P;
        bool roundb = false, sticky = false;
        if (shift > 0) {
            const int rb = shift - 1, qr = rb >> 5, br2 = rb & 31;
            const uint32_t wr = word(qr);
            roundb = (wr >> br2) & 1u;
            sticky = low < qr || (br2 && (wr & ((1u << br2) - 1u)) != 0);
        }
        // mantissa words: bits [shift, shift + P) of the magnitude
        uint32_t mant[CRT_MAXW];
        {
            const int q = shift >> 5, rs = shift & 31;  // arithmetic: floor for negative shifts
#pragma unroll
            for (int w = 0; w < CRT_MAXW; w++) {
                if (w < 2 * n) {
                    const uint32_t lo = word(q + w), hi = word(q + w + 1);
                    mant[w] = __builtin_amdgcn_alignbit(hi, lo, rs);
                } else {
                    mant[w] = 0;
                }
            }
        }
        if (roundb && (sticky || (mant[0] & 1u))) {
            uint32_t c = 1;
#pragma unroll
            for (int w = 0; w < CRT_MAXW; w++)
                if (w < 2 * n) {
                    mant[w] += c;
                    c = c && mant[w] == 0;
                }
            if (c) {
#pragma unroll
                for (int w = 0; w < CRT_MAXW; w++)
                    if (w < 2 * n) mant[w] = (w == 2 * n - 1) ? 0x80000000u : 0u;
                shift += 1;
            }
        }
        const int64_t expo = (int64_t)scan_a[i].x + scan_b[j].x + shift + P;
        if (expo > (int64_t)((1 << 30) - 1)) {
            if constexpr (NL > 0) {
                if (srec) {
                    uint32_t zw[2 * (NL > 0 ? NL : 1)] = {};
                    crt_sub_store<NL>(srec, zw, negative ? -KIND_INF : KIND_INF, 0);
                    continue;
                }
            }
            for (int w = 0; w < 2 * n; w++) rec[w] = 0;
            rec[2 * n] = (uint32_t)(negative ? -KIND_INF : KIND_INF);
            rec[2 * n + 1] = 0;
            continue;
        }
        if (expo < (int64_t)(1 - (1 << 30))) {
            if (srec) continue;  // the entry rounds to zero: RN(s - 0) = s
            for (int w = 0; w < 2 * n; w++) rec[w] = 0;
            rec[2 * n] = KIND_ZERO;
            rec[2 * n + 1] = 0;
            continue;
        }
        if constexpr (NL > 0) {
            if (srec) {
                crt_sub_store<NL>(srec, mant, negative ? -KIND_REGULAR : KIND_REGULAR, (int32_t)expo);
                continue;
            }
        }
#pragma unroll
        for (int w = 0; w < CRT_MAXW; w++)
            if (w < 2 * n) rec[w] = mant[w];
        rec[2 * n] = (uint32_t)(negative ? -KIND_REGULAR : KIND_REGULAR);
        rec[2 * n + 1] = (uint32_t)(int32_t)expo;
    }
}

__global__ void zero_values(uint32_t *__restrict__ out, int n, int64_t m, int64_t nn,
                            int upper, int packed) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * nn) return;
    int64_t i = idx % m, j = idx / m;
    if (upper && i > j) return;
    int64_t o = packed ? (j * (j + 1) / 2 + i) : (i + j * m);
    uint32_t *rec = out + (size_t)o * (2 * n + 2);
    for (int w = 0; w < 2 * n; w++) rec[w] = 0;
    rec[2 * n] = KIND_ZERO;
    rec[2 * n + 1] = 0;
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
namespace {

enum Slot {
    S_SCAN_A = 0, S_SCAN_B, S_PLANES_A, S_PLANES_B, S_X, S_POW_A, S_POW_B, S_MODS, S_CRT,
    S_TILES, S_ENC_A, S_ENC_B, S_CRT2, S_BATCH_SCAN, S_COUNT
};

struct Operand {
    const uint32_t *data;
    int64_t V, K, sv, sp;
};

void launch_scan(const Operand &op, int n, int4 *out, hipStream_t s) {
    if (op.V == 0) return;
    if (op.sv == 1 && op.sp != 1) {
        // position chunks of at least 32 (8 records a thread), at most 64 chunks
        const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(64, op.K / 32));
        const int64_t pch = (op.K + chunks - 1) / chunks;
        hipLaunchKernelGGL(scan_init, dim3((unsigned)((op.V + 255) / 256)), dim3(256), 0, s, op.V, out);
        hipLaunchKernelGGL(scan_vec_fast, dim3((unsigned)((op.V + 63) / 64), (unsigned)chunks), dim3(256), 0, s,
                           op.data, n, op.V, op.K, op.sv, op.sp, pch, out);
    } else {
        hipLaunchKernelGGL(scan_pos_fast, dim3((unsigned)((op.V + 3) / 4)), dim3(256), 0, s,
                           op.data, n, op.V, op.K, op.sv, op.sp, out);
    }
    HIP_CHECK(hipGetLastError());
}

struct Window {
    int64_t width = 0;  // max_v (emax_v - lo_v)
    int64_t emin = INT64_MAX, emax = INT64_MIN;
    bool nonfinite = false, any = false;
};

// The operand's window from its per-vector scan (the largest top-to-lowest
// bit spread of a vector, the exponent range, the flags), written to
// out[0..5) = width, emin, emax, non-finite, any: one block, so the host
// reads five words instead of the whole scan.
__global__ __launch_bounds__(256) void scan_summary(const int4 *__restrict__ scan, int64_t V, int64_t *out) {
    __shared__ int64_t red[4][4];
    int64_t width = 0, emin = INT64_MAX, emax = INT64_MIN;
    int flags = 0;
    for (int64_t v = threadIdx.x; v < V; v += 256) {
        const int4 x = scan[v];
        flags |= x.w;
        if (x.w & 2) {
            width = max(width, (int64_t)x.z - x.x);
            emin = min(emin, (int64_t)x.y);
            emax = max(emax, (int64_t)x.z);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        width = max(width, (int64_t)__shfl_xor(width, off, 64));
        emin = min(emin, (int64_t)__shfl_xor(emin, off, 64));
        emax = max(emax, (int64_t)__shfl_xor(emax, off, 64));
        flags |= __shfl_xor(flags, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = width;
        red[wave][1] = emin;
        red[wave][2] = emax;
        red[wave][3] = flags;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            width = max(width, red[w][0]);
            emin = min(emin, red[w][1]);
            emax = max(emax, red[w][2]);
            flags |= (int)red[w][3];
        }
        out[0] = width;
        out[1] = emin;
        out[2] = emax;
        out[3] = flags & 1;
        out[4] = (flags >> 1) & 1;
    }
}

Window window_of(const int64_t *w) {
    Window r;
    r.width = w[0];
    r.emin = w[1];
    r.emax = w[2];
    r.nonfinite = w[3] != 0;
    r.any = w[4] != 0;
    return r;
}


template <int TJMAX>
void launch_encode_t(const Operand &op, int n, const int4 *scan, int TJ, const double *pow2,
                     const uint32_t *mq, const double *mqinv, int L, int single, int8_t *planes,
                     int64_t plane_stride, int64_t k_pad, int64_t v_pad, hipStream_t s) {
    size_t lds = 1024 * (2 * (size_t)n + 2) * 4;
    if (lds > 65536)
        HIP_CHECK(hipFuncSetAttribute((const void *)encode_kernel<TJMAX>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    dim3 grid((unsigned)(k_pad / 32), (unsigned)(v_pad / 32));
    hipLaunchKernelGGL(encode_kernel<TJMAX>, grid, dim3(256), lds, s, op.data, n, op.V, op.K,
                       op.sv, op.sp, scan, TJ, pow2, mq, mqinv, L, single, planes, plane_stride,
                       k_pad);
    HIP_CHECK(hipGetLastError());
}

void launch_encode(const Operand &op, int n, const int4 *scan, int TJ, const double *pow2,
                   const uint32_t *mq, const double *mqinv, int L, int single, int8_t *planes,
                   int64_t plane_stride, int64_t k_pad, int64_t v_pad, hipStream_t s) {
    if (TJ <= 16)
        launch_encode_t<16>(op, n, scan, TJ, pow2, mq, mqinv, L, single, planes, plane_stride,
                            k_pad, v_pad, s);
    else if (TJ <= 32)
        launch_encode_t<32>(op, n, scan, TJ, pow2, mq, mqinv, L, single, planes, plane_stride,
                            k_pad, v_pad, s);
    else
        launch_encode_t<64>(op, n, scan, TJ, pow2, mq, mqinv, L, single, planes, plane_stride,
                            k_pad, v_pad, s);
}

// Tables of the matrix-core encode for inner byte count 32 nk: th, tl
// (count x 32 nk signed bytes), E_l, q_l, floor(2^32 / q_l).  Laid out in one
// buffer: th | tl | E | q | mbar.
struct EncodeTables {
    std::vector<uint8_t> bytes;
    size_t th = 0, tl = 0, e = 0, q = 0, mbar = 0;
};

EncodeTables encode_tables(const RnsPlan &plan, int nk) {
    const int L = plan.count, bpad = 32 * nk;
    EncodeTables t;
    t.th = 0;
    t.tl = (size_t)L * bpad;
    t.e = 2 * (size_t)L * bpad;
    t.q = t.e + 4 * (size_t)L;
    t.mbar = t.q + 4 * (size_t)L;
    t.bytes.assign(t.mbar + 4 * (size_t)L, 0);
    for (int l = 0; l < L; l++) {
        const int64_t q = plan.q[l];
        int64_t pw = 1 % q, sum = 0;
        for (int b = 0; b < bpad; b++) {
            const int64_t sym = pw >= (q + 1) / 2 ? pw - q : pw;
            const int64_t lo = ((sym + 128) & 255) - 128, hi = (sym - lo) / 256;
            t.bytes[t.th + (size_t)l * bpad + b] = (uint8_t)(int8_t)hi;
            t.bytes[t.tl + (size_t)l * bpad + b] = (uint8_t)(int8_t)lo;
            sum += sym;
            pw = pw * 256 % q;
        }
        const int64_t c = 128 * sum, base = int64_t(1) << 30;
        const uint32_t e = (uint32_t)(base + (((c - base) % q) + q) % q);
        const uint32_t qq = (uint32_t)q, mb = (uint32_t)((uint64_t(1) << 32) / (uint64_t)q);
        std::memcpy(&t.bytes[t.e + 4 * (size_t)l], &e, 4);
        std::memcpy(&t.bytes[t.q + 4 * (size_t)l], &qq, 4);
        std::memcpy(&t.bytes[t.mbar + 4 * (size_t)l], &mb, 4);
    }
    return t;
}

// Inner bytes the matrix-core encode handles: 3 MFMA k-steps (768 bits)
// keep every partial sum inside 31 bits.
constexpr int ENC_MAX_NK = 3;

void launch_encode_mfma(const Operand &op, int n, const int4 *scan, int nk, const uint8_t *d_tab,
                        const EncodeTables &t, int L, int single, int8_t *planes,
                        int64_t plane_stride, int64_t k_pad, int64_t v_pad, hipStream_t s) {
    const int nblocks = (L + 31) / 32, trow = 32 * nk + 16;
    const int64_t tiles_v = v_pad / ENC_TV, tiles_p = k_pad / ENC_TP;
    const int64_t tiles = tiles_v * tiles_p;
    // few tiles: the modulus blocks side by side (the grid would not fill the chip)
    const int gy = tiles < 1024 ? nblocks : 1;
    const size_t lds = 2 * (size_t)(gy > 1 ? 32 : 32 * nblocks) * trow + 64 * (size_t)ENC_ROW;
    if (lds > 65536)
        HIP_CHECK(hipFuncSetAttribute((const void *)encode_mfma_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, 2048);
    hipLaunchKernelGGL(encode_mfma_kernel, dim3(grid, gy), dim3(256), lds, s, op.data, n, op.V, op.K,
                       op.sv, op.sp, scan, nk, (const int8_t *)(d_tab + t.th),
                       (const int8_t *)(d_tab + t.tl), (const uint32_t *)(d_tab + t.e),
                       (const uint32_t *)(d_tab + t.q), (const uint32_t *)(d_tab + t.mbar), L,
                       single, planes, plane_stride, k_pad, tiles_v, tiles_p);
    HIP_CHECK(hipGetLastError());
}

// 2^(24 j) mod q_l for j < TJ, as doubles (count x TJ).
std::vector<double> pow2_table(const RnsPlan &plan, int TJ) {
    std::vector<double> t((size_t)plan.count * TJ);
    for (int l = 0; l < plan.count; l++) {
        uint64_t q = plan.q[l], v = 1 % q, step = (1u << 24) % q;
        for (int j = 0; j < TJ; j++) {
            t[(size_t)l * TJ + j] = (double)v;
            v = v * step % q;
        }
    }
    return t;
}

thread_local RnsStats t_stats;
bool g_profile = false;
bool g_encode_mfma = true;  // matrix-core encode where the window allows
bool g_crt_mfma = true;     // matrix-core CRT where the plan allows

// Events bracketing the four stages (only when profiling).
struct StageEvents {
    hipEvent_t ev[5] = {};
    bool on;
    explicit StageEvents(bool enable) : on(enable) {
        if (on)
            for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
    }
    void mark(int i, hipStream_t s) {
        if (on) HIP_CHECK(hipEventRecord(ev[i], s));
    }
    void collect(RnsStats &st) {
        if (!on) return;
        HIP_CHECK(hipEventSynchronize(ev[4]));
        for (int i = 0; i < 4; i++) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            st.stage_ms[i] = ms;
        }
    }
    ~StageEvents() {
        if (on)
            for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

unsigned padded_grid(int64_t total) { return (unsigned)((total + 7) / 8 * 8); }

// Product kernel selection (measured on MI355X, 4096^3 at 256 bits): one-digit
// moduli use product_kernel<1> (128 x 128 tiles, three workgroups per CU),
// two-digit moduli product2_kernel<2, 128> (double-buffered).
// SDPX_GPU_PRODUCT_KERNEL=1 selects the first form for both, =2 the
// double-buffered form for both (256 x 128 one-digit tiles), for comparison.
int product_kernel_choice() {
    static const int choice = [] {
        const char *e = std::getenv("SDPX_GPU_PRODUCT_KERNEL");
        return e && e[0] == '1' ? 1 : (e && e[0] == '2' ? 2 : 0);
    }();
    return choice;
}

// Tile order of the product kernels.  Each XCD runs a contiguous range of
// the list with some 64 workgroups in flight, and they read the same k
// slices of their operand panels at about the same time; listing tiles in
// bands of `group` column tiles (rows outer, columns inner inside a band)
// makes that window a compact block of about group x (64 / group) tiles, so
// each panel slice is fetched into the XCD's L2 once per block instead of
// once per tile.  SDPX_GPU_TILE_GROUP overrides the band width (0: one band).
int tile_group() {
    static const int g = [] {
        const char *e = std::getenv("SDPX_GPU_TILE_GROUP");
        return e ? std::atoi(e) : 8;
    }();
    return g;
}

// (row block, column tile) pairs in band order; keep(r, c) filters (upper
// products keep the blocks that meet the upper triangle).
template <typename Keep>
std::vector<int2> banded_tiles(int rows, int cols, Keep keep) {
    const int g = tile_group() > 0 ? tile_group() : cols;
    std::vector<int2> out;
    for (int c0 = 0; c0 < cols; c0 += g)
        for (int r = 0; r < rows; r++)
            for (int c = c0; c < std::min(cols, c0 + g); c++)
                if (keep(r, c)) out.push_back(make_int2(r, c));
    return out;
}

template <int D, int BM>
void launch_product2(const int8_t *pla, const int8_t *plb, int64_t a_plane, int64_t b_plane,
                     int64_t k_pad, int64_t k0, int64_t k1, const uint32_t *d_q, const uint32_t *d_inv,
                     const double *d_qinv, int single, int mod_first, int mod_count, const int2 *tiles,
                     int ntiles, int tiles_m, int tiles_n, int acc, uint16_t *X, hipStream_t s) {
    constexpr int STAGE = D * (BM + TILE) * LDS_ROW;
    constexpr int XT = TILE * (BM + 2) * 2;
    constexpr size_t lds = 2 * STAGE > XT ? 2 * STAGE : XT;
    auto kern = product2_kernel<D, BM>;
    static bool attr = false;
    if (!attr && lds > 65536) {
        HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
        attr = true;
    }
    hipLaunchKernelGGL(kern, dim3(padded_grid((int64_t)mod_count * ntiles)), dim3(256), lds, s, pla, plb,
                       a_plane, b_plane, k_pad, k0, k1, d_q, d_inv, d_qinv, single, mod_first, mod_count,
                       tiles, ntiles, tiles_m, tiles_n, acc, X);
    HIP_CHECK(hipGetLastError());
}

}  // namespace

const RnsStats &rns_last_stats() { return t_stats; }
thread_local RnsTotals t_totals;
RnsTotals &rns_totals() { return t_totals; }
void rns_set_profiling(bool on) { g_profile = on; }
bool rns_profiling() { return g_profile; }

// Device copies of the tables that depend only on the modulus plan (and the
// encode window) or on the tile grid.  They are made once per device and
// kept for the life of the process, so a product uploads nothing but reads
// back its two scan windows.
struct PlanTables {
    uint32_t *q = nullptr, *inv = nullptr;
    double *qinv = nullptr, *crt = nullptr;
    uint8_t *crt2 = nullptr;
    EncodeTables enc[ENC_MAX_NK + 1];
    uint8_t *d_enc[ENC_MAX_NK + 1] = {};
    std::map<int, double *> pow2;
};
std::mutex g_tables_mu;
std::map<std::pair<int, const RnsPlan *>, PlanTables> g_plan_tables;
std::map<std::vector<int>, int2 *> g_tile_tables;

template <typename T>
T *upload_once(const void *src, size_t bytes) {
    void *d = nullptr;
    HIP_CHECK(hipMalloc(&d, std::max<size_t>(bytes, 16)));
    if (bytes) HIP_CHECK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
    return (T *)d;
}

int current_device() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    return dev;
}

// The plan's moduli, CRT and matrix-core CRT tables (made on first use).
PlanTables &plan_tables(const RnsPlan &plan) {
    std::lock_guard<std::mutex> g(g_tables_mu);
    PlanTables &t = g_plan_tables[{current_device(), &plan}];
    if (!t.q) {
        const size_t L = plan.count, J = plan.limbs;
        t.q = upload_once<uint32_t>(plan.q.data(), 4 * L);
        t.inv = upload_once<uint32_t>(plan.inv.data(), 4 * L);
        t.qinv = upload_once<double>(plan.qinv.data(), 8 * L);
        std::vector<double> crt(plan.c.begin(), plan.c.begin() + L * J);
        crt.insert(crt.end(), plan.m.begin(), plan.m.begin() + J);
        t.crt = upload_once<double>(crt.data(), 8 * crt.size());
        std::vector<uint8_t> c2(plan.crt_table.begin(), plan.crt_table.end());
        const uint8_t *m32 = (const uint8_t *)plan.m32.data();
        c2.insert(c2.end(), m32, m32 + 4 * plan.m32.size());
        if (plan.m32.size() < CRT_NU_MAX) c2.resize(c2.size() + 4 * (CRT_NU_MAX - plan.m32.size()), 0);  // M zero-padded
        t.crt2 = upload_once<uint8_t>(c2.data(), c2.size());
    }
    return t;
}

// Matrix-core encode tables for nk k-steps of value bytes.
const EncodeTables &encode_tables_dev(const RnsPlan &plan, int nk, uint8_t *&dev) {
    PlanTables &t = plan_tables(plan);
    std::lock_guard<std::mutex> g(g_tables_mu);
    if (!t.d_enc[nk]) {
        t.enc[nk] = encode_tables(plan, nk);
        t.d_enc[nk] = upload_once<uint8_t>(t.enc[nk].bytes.data(), t.enc[nk].bytes.size());
    }
    dev = t.d_enc[nk];
    return t.enc[nk];
}

// 24-bit chunk powers for the VALU encode form.
double *pow2_table_dev(const RnsPlan &plan, int TJ) {
    PlanTables &t = plan_tables(plan);
    std::lock_guard<std::mutex> g(g_tables_mu);
    double *&d = t.pow2[TJ];
    if (!d) {
        const std::vector<double> h = pow2_table(plan, TJ);
        d = upload_once<double>(h.data(), 8 * h.size());
    }
    return d;
}

// The product kernels' tile lists (tl then tl2) for one grid.
int2 *tile_tables_dev(const std::vector<int2> &tl, const std::vector<int2> &tl2, const std::vector<int> &key) {
    std::lock_guard<std::mutex> g(g_tables_mu);
    std::vector<int> k = key;
    k.push_back(current_device());
    int2 *&d = g_tile_tables[k];
    if (!d) {
        std::vector<int2> all(tl);
        all.insert(all.end(), tl2.begin(), tl2.end());
        d = upload_once<int2>(all.data(), sizeof(int2) * all.size());
    }
    return d;
}

// The fixed device memory of the allocation-free exact product
// (sdpx_gpu_rns_gemm_dev_noalloc): one array per device, held until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: it uses the device address looked up once per device at set-up.
__device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> address of that device's scratch
}  // namespace

void rns_noalloc_setup() {
    const int dev = current_device();
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr.count(dev)) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        throw Error(SDPX_GPU_DEVICE_ERROR, std::string("scratch address lookup failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
}

void *rns_noalloc_scratch() {
    const int dev = current_device();
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end())
        throw Error(SDPX_GPU_DEVICE_ERROR, "allocation-free scratch not set up on this device (sdpx_gpu_init)");
    return it->second;
}

namespace {
// Device buffers of one exact product: workspace slots, or (allocation-free
// form) consecutive 256-byte-aligned pieces of the caller's fixed scratch,
// whose overflow is a capacity error raised before any output is written.
struct Buffers {
    Workspace *ws;
    uint8_t *base;
    size_t cap, top = 0;
    bool fixed() const { return base != nullptr; }
    void *get(int slot, size_t bytes) {
        if (!base) return ws->get(slot, bytes);
        bytes = (bytes + 255) & ~size_t(255);
        if (top + bytes > cap)
            throw Error(SDPX_GPU_CAPACITY, "exact product temporaries need " + std::to_string((top + bytes) >> 20) +
                                               "+ MiB, above the " + std::to_string(cap >> 20) +
                                               " MiB allocation-free scratch");
        void *p = base + top;
        top += bytes;
        return p;
    }
};
}  // namespace

// Whether a plan's CRT runs on the matrix cores (crt_mfma_kernel): the MFMA
// form holds every k-slice's fragments, at most 2 per 32-row table block.
// The matrix-core CRT's compiled width for a plan: words of M (NUMAX) and
// 32-row table blocks (NBMAX); 0 when none fits.
static int crt_numax(const RnsPlan &plan) {
    const int nu = plan.crt_nu, nbl = plan.crt_npad / 32;
    return nu <= 20 && nbl <= 3 ? 20 : (nu <= 24 && nbl <= 3 ? 24 : (nu <= 32 && nbl <= 4 ? 32 : (nu <= 40 && nbl <= 5 ? 40 : 0)));
}
static int crt_nbmax(int numax) { return numax <= 24 ? 3 : (numax == 32 ? 4 : 5); }

bool crt_matrix_core(const RnsPlan &plan, int n) {
    const int numax = crt_numax(plan);
    return g_crt_mfma && numax > 0 && n <= 8 && plan.crt_lpad / 16 <= 2 * crt_nbmax(numax);
}

// Whether the matrix-core CRT's buffer offsets fit this output shape (tiles
// padded to at most 256 rows or columns): the lane offset (8 planes and a
// tile column) and the plane offset (up to 7 planes) each below 2^31, so
// their sum stays below 2^32.
static bool crt_offsets_fit(int64_t m, int64_t nn) {
    const uint64_t tiles = (uint64_t)((m + 2 * TILE - 1) / TILE) * (uint64_t)((nn + 2 * TILE - 1) / TILE);
    return 8 * tiles * TILE * TILE * sizeof(uint16_t) + 4 * TILE * TILE < (1ull << 31);
}

int rns_gemm_device(const RnsGemmArgs &g) {
    Device &dev = device();
    hipStream_t s = dev.stream;
    Buffers ws{&dev.ws, (uint8_t *)g.scratch, g.scratch_bytes};
    const bool fixed = ws.fixed();
    // allocation-free form: host tables copied into the scratch stay alive
    // here until the stream has run the copies
    std::vector<std::vector<uint8_t>> keep;
    auto table_copy = [&](const void *src, size_t bytes) -> void * {
        void *d = ws.get(-1, std::max<size_t>(bytes, 16));
        keep.emplace_back((const uint8_t *)src, (const uint8_t *)src + bytes);
        if (bytes) HIP_CHECK(hipMemcpyAsync(d, keep.back().data(), bytes, hipMemcpyHostToDevice, s));
        return d;
    };
    struct Drain {  // the copies above read host memory owned by this call
        bool on;
        hipStream_t s;
        ~Drain() {
            if (on) (void)hipStreamSynchronize(s);
        }
    } drain{fixed, s};
    const int n = g.limbs;
    const bool tan = (g.ta == 'N' || g.ta == 'n');
    const bool tbn = (g.tb == 'N' || g.tb == 'n');
    const int64_t m = g.m, nn = g.n, k = g.k;
    const bool packed = g.upper && m == nn && g.out_len == (uint64_t)(nn * (nn + 1) / 2);
    t_stats = RnsStats{};
    if (g.sub_done) *g.sub_done = 0;
    if (m == 0 || nn == 0) return SDPX_GPU_OK;

    Operand oa{(const uint32_t *)g.a, m, k, tan ? 1 : g.lda, tan ? g.lda : 1};
    Operand ob{(const uint32_t *)g.b, nn, k, tbn ? g.ldb : 1, tbn ? 1 : g.ldb};
    const bool same = g.a == g.b && oa.sv == ob.sv && oa.sp == ob.sp && m == nn && g.lda == g.ldb;

    auto zero_fill = [&]() {
        int64_t total = m * nn;
        hipLaunchKernelGGL(zero_values, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           (uint32_t *)g.out, n, m, nn, g.upper ? 1 : 0, packed ? 1 : 0);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s));
    };
    if (k == 0) {
        if (g.split_capacity > 0 && g.split_used) *g.split_used = 1;
        zero_fill();
        return SDPX_GPU_OK;
    }

    // 1. scan
    StageEvents sev(g_profile && !fixed);
    sev.mark(0, s);
    int4 *scan_a, *scan_b;
    int64_t rb_host[16] = {};
    int64_t *rb;
    if (g.summary) {  // batched form: scanned and waited for by rns_gemm_batch
        scan_a = (int4 *)g.scan_a_in;
        scan_b = same ? scan_a : (int4 *)g.scan_b_in;
        rb = (int64_t *)g.summary;
    } else {
    scan_a = (int4 *)ws.get(S_SCAN_A, sizeof(int4) * m);
    scan_b = same ? scan_a : (int4 *)ws.get(S_SCAN_B, sizeof(int4) * nn);
    launch_scan(oa, n, scan_a, s);
    if (!same) launch_scan(ob, n, scan_b, s);
    // the five-word summaries: pinned read-back words, or (allocation-free
    // form) 16 words of the scratch copied back
    rb = fixed ? (int64_t *)ws.get(-1, sizeof rb_host) : readback_words();
    hipLaunchKernelGGL(scan_summary, dim3(1), dim3(256), 0, s, (const int4 *)scan_a, m, rb);
    if (!same) hipLaunchKernelGGL(scan_summary, dim3(1), dim3(256), 0, s, (const int4 *)scan_b, nn, rb + 8);
    HIP_CHECK(hipGetLastError());
    if (fixed) {
        HIP_CHECK(hipMemcpyAsync(rb_host, rb, sizeof rb_host, hipMemcpyDeviceToHost, s));
        rb = rb_host;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    }
    Window wa = window_of(rb), wb = same ? wa : window_of(rb + 8);
    if (wa.nonfinite || wb.nonfinite) return SDPX_GPU_DECLINED;
    if ((wa.any && wa.emax - wa.emin > MAX_SPREAD) || (wb.any && wb.emax - wb.emin > MAX_SPREAD))
        return SDPX_GPU_DECLINED;
    if (!wa.any || !wb.any) {
        if (g.split_capacity > 0 && g.split_used) *g.split_used = 1;
        zero_fill();
        return SDPX_GPU_OK;
    }
    const int TJa = (int)((wa.width + 23) / 24) + 1, TJb = (int)((wb.width + 23) / 24) + 1;
    if (TJa > 64 || TJb > 64) return SDPX_GPU_DECLINED;
    const RnsPlan *plan = rns_plan_for_bits(rns_needed_bits(wa.width, wb.width, k));
    if (!plan) return SDPX_GPU_DECLINED;
    if (plan->limbs > 128) return SDPX_GPU_DECLINED;
    const int L = plan->count, G = plan->planes, single = plan->single;
    // expansion output: enough P-bit chunks for the largest exact value
    int split = 0;
    if (g.split_capacity > 0) {
        split = (int)std::ceil((plan->bits + 2.0) / (64.0 * n)) + 1;
        if (split > g.split_capacity || !crt_matrix_core(*plan, n) || !crt_offsets_fit(m, nn)) return SDPX_GPU_DECLINED;
        if (g.split_used) *g.split_used = split;
    }
    t_stats.moduli = L;
    t_stats.planes = G;
    t_stats.bits = plan->bits;
    t_stats.width_a = wa.width;
    t_stats.width_b = wb.width;

    // tables: moduli, inverses, 1/q, pow2 (A, B), CRT
    const int64_t k_pad = (k + KSTEP - 1) / KSTEP * KSTEP;
    // the second product kernel takes 256-row blocks of op(A): pad both
    // sides to 256 (equal strides keep the Gram form's shared planes valid)
    const bool p2 = product_kernel_choice() == 2;
    const int64_t PADM = p2 ? 2 * TILE : TILE;
    const int64_t m_pad = (m + PADM - 1) / PADM * PADM, n_pad = (nn + PADM - 1) / PADM * PADM;
    const int tiles_m = (int)(m_pad / TILE), tiles_n = (int)(n_pad / TILE);
    const std::vector<int2> tl =
        banded_tiles(tiles_m, tiles_n, [&](int tm, int tn) { return !g.upper || tm <= tn; });
    const int ntiles = (int)tl.size();

    const int J = plan->limbs;
    // encode tables: the matrix-core form when the window fits 3 k-steps,
    // else 24-bit chunk powers for the VALU form
    const int nka = (int)(((wa.width + 7) / 8 + 31) / 32), nkb = (int)(((wb.width + 7) / 8 + 31) / 32);
    const bool mfma_a = g_encode_mfma && nka <= ENC_MAX_NK, mfma_b = g_encode_mfma && nkb <= ENC_MAX_NK;
    static const EncodeTables no_tables;
    uint8_t *d_eta = nullptr, *d_etb = nullptr;
    uint32_t *d_q, *d_inv;
    double *d_qinv, *d_crt, *d_pa = nullptr, *d_pb = nullptr;
    const uint8_t *d_ct;
    EncodeTables eta_fixed, etb_fixed;  // allocation-free form: made for this call
    const EncodeTables *eta = &no_tables, *etb = &no_tables;
    if (!fixed) {
        PlanTables &ptab = plan_tables(*plan);
        d_q = ptab.q;
        d_inv = ptab.inv;
        d_qinv = ptab.qinv;
        d_crt = ptab.crt;
        d_ct = ptab.crt2;
        if (mfma_a) eta = &encode_tables_dev(*plan, nka, d_eta);
        if (!same && mfma_b) etb = &encode_tables_dev(*plan, nkb, d_etb);
        if (!mfma_a) d_pa = pow2_table_dev(*plan, TJa);
        if (!same && !mfma_b) d_pb = pow2_table_dev(*plan, TJb);
    } else {
        d_q = (uint32_t *)table_copy(plan->q.data(), 4 * (size_t)L);
        d_inv = (uint32_t *)table_copy(plan->inv.data(), 4 * (size_t)L);
        d_qinv = (double *)table_copy(plan->qinv.data(), 8 * (size_t)L);
        std::vector<double> crt(plan->c.begin(), plan->c.begin() + (size_t)L * J);
        crt.insert(crt.end(), plan->m.begin(), plan->m.begin() + J);
        d_crt = (double *)table_copy(crt.data(), 8 * crt.size());
        std::vector<uint8_t> c2(plan->crt_table.begin(), plan->crt_table.end());
        const uint8_t *m32 = (const uint8_t *)plan->m32.data();
        c2.insert(c2.end(), m32, m32 + 4 * plan->m32.size());
        if (plan->m32.size() < CRT_NU_MAX) c2.resize(c2.size() + 4 * (CRT_NU_MAX - plan->m32.size()), 0);
        d_ct = (const uint8_t *)table_copy(c2.data(), c2.size());
        if (mfma_a) {
            eta_fixed = encode_tables(*plan, nka);
            d_eta = (uint8_t *)table_copy(eta_fixed.bytes.data(), eta_fixed.bytes.size());
            eta = &eta_fixed;
        } else {
            const std::vector<double> h = pow2_table(*plan, TJa);
            d_pa = (double *)table_copy(h.data(), 8 * h.size());
        }
        if (!same && mfma_b) {
            etb_fixed = encode_tables(*plan, nkb);
            d_etb = (uint8_t *)table_copy(etb_fixed.bytes.data(), etb_fixed.bytes.size());
            etb = &etb_fixed;
        } else if (!same) {
            const std::vector<double> h = pow2_table(*plan, TJb);
            d_pb = (double *)table_copy(h.data(), 8 * h.size());
        }
    }
    // 256 x 128 blocks of the second product kernel (one-digit moduli)
    std::vector<int2> tl2;
    if (p2)
        tl2 = banded_tiles(tiles_m / 2, tiles_n, [&](int tm2, int tn) {
            return !g.upper || (int64_t)tm2 * 2 * TILE <= (int64_t)tn * TILE + TILE - 1;
        });
    const int ntiles2 = (int)tl2.size();
    int2 *d_tiles;
    if (!fixed) {
        d_tiles = tile_tables_dev(tl, tl2, {tiles_m, tiles_n, g.upper ? 1 : 0, p2 ? 1 : 0, tile_group()});
    } else {
        std::vector<int2> all(tl);
        all.insert(all.end(), tl2.begin(), tl2.end());
        d_tiles = (int2 *)table_copy(all.data(), sizeof(int2) * all.size());
    }
    int2 *d_tiles2 = d_tiles + ntiles;

    // every buffer is in place before the first encode launch (in the
    // allocation-free form a capacity error is raised here, output untouched)
    const int64_t a_plane = m_pad * k_pad, b_plane = n_pad * k_pad;
    int8_t *pla = (int8_t *)ws.get(S_PLANES_A, (size_t)G * a_plane);
    int8_t *plb = same ? pla : (int8_t *)ws.get(S_PLANES_B, (size_t)G * b_plane);
    // planes up to the matrix-core CRT's padded count (read there, times zero coefficients)
    uint16_t *X = (uint16_t *)ws.get(S_X, (size_t)std::max(L, plan->crt_lpad) * tiles_m * tiles_n * TILE * TILE * 2);

    // 2. encode
    sev.mark(1, s);
    if (mfma_a)
        launch_encode_mfma(oa, n, scan_a, nka, d_eta, *eta, L, single, pla, a_plane, k_pad, m_pad, s);
    else
        launch_encode(oa, n, scan_a, TJa, d_pa, d_q, d_qinv, L, single, pla, a_plane, k_pad, m_pad, s);
    if (!same) {
        if (mfma_b)
            launch_encode_mfma(ob, n, scan_b, nkb, d_etb, *etb, L, single, plb, b_plane, k_pad, n_pad, s);
        else
            launch_encode(ob, n, scan_b, TJb, d_pb, d_q, d_qinv, L, single, plb, b_plane, k_pad,
                          n_pad, s);
    }

    // 3. products, in exact INT32 chunks of the inner dimension
    sev.mark(2, s);
    for (int64_t k0 = 0; k0 < k_pad; k0 += KCHUNK) {
        const int64_t k1 = std::min<int64_t>(k_pad, k0 + KCHUNK);
        const int acc = k0 > 0 ? 1 : 0;
        if (p2 && single > 0)
            launch_product2<1, 2 * TILE>(pla, plb, a_plane, b_plane, k_pad, k0, k1, d_q, d_inv, d_qinv,
                                         single, 0, single, d_tiles2, ntiles2, tiles_m, tiles_n, acc,
                                         X, s);
        if (product_kernel_choice() != 1 && L > single)
            launch_product2<2, TILE>(pla, plb, a_plane, b_plane, k_pad, k0, k1, d_q, d_inv, d_qinv,
                                     single, single, L - single, d_tiles, ntiles, tiles_m, tiles_n, acc,
                                     X, s);
        if (!p2 && single > 0) {
            hipLaunchKernelGGL(product_kernel<1>, dim3(padded_grid(single * ntiles)), dim3(256), 0, s,
                               pla, plb, a_plane, b_plane, k_pad, k0, k1, d_q, d_inv, d_qinv,
                               single, 0, single, d_tiles, ntiles, tiles_m, tiles_n, acc, X);
            HIP_CHECK(hipGetLastError());
        }
        if (product_kernel_choice() == 1 && L > single) {
            hipLaunchKernelGGL(product_kernel<2>, dim3(padded_grid((L - single) * ntiles)),
                               dim3(256), 0, s, pla, plb, a_plane, b_plane, k_pad, k0, k1, d_q,
                               d_inv, d_qinv, single, single, L - single, d_tiles, ntiles, tiles_m,
                               tiles_n, acc, X);
            HIP_CHECK(hipGetLastError());
        }
    }

    // 4. crt + rounding
    sev.mark(3, s);
    t_stats.product_launches = ((k_pad + KCHUNK - 1) / KCHUNK) * ((single > 0) + (L > single));
    const double *d_c = d_crt, *d_m = d_crt + (size_t)L * J;
    const int up = g.upper ? 1 : 0, pk = packed ? 1 : 0;
    uint32_t *outw = (uint32_t *)g.out;
    const bool crt_mma = crt_matrix_core(*plan, n) && crt_offsets_fit(m, nn);
    // the fused subtraction needs the matrix-core CRT at a compiled width
    const bool fuse_sub = g.sub && crt_mma && split == 0 && g.upper && m == nn && (n == 2 || n == 4 || n == 8);
    if (g.sub_done) *g.sub_done = fuse_sub ? 1 : 0;
    if (crt_mma) {
        const int lpad = plan->crt_lpad, npad = plan->crt_npad, nu = plan->crt_nu;
        const size_t tb = plan->crt_table.size();
        const int trow = 2 * lpad + 16;
        const int numax = crt_numax(*plan);
        // the table, then each thread's column of words
        const size_t lds = (size_t)npad * trow + (size_t)numax * 128 * 4;
        const int64_t items = (int64_t)ntiles * TILE;
        const unsigned grid = (unsigned)std::min<int64_t>(items, 4096);
        auto launch = [&](auto kern) {
            if (lds > 65536)
                HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds));
            hipLaunchKernelGGL(kern, dim3(grid), dim3(128), lds, s, (const uint16_t *)X, L, lpad,
                               (const int2 *)d_tiles, ntiles, tiles_m, tiles_n, (const int8_t *)d_ct,
                               plan->crt_nt, npad, (const uint32_t *)(d_ct + tb), nu, plan->m_top,
                               (const int4 *)scan_a, (const int4 *)scan_b, m, nn, up, pk, n, outw, split,
                               (int64_t)(m * nn), fuse_sub ? (uint32_t *)g.sub : nullptr, g.sub_ld);
        };
        auto pick = [&](auto k2, auto k4, auto k8, auto kn) {
            if (n == 2) launch(k2);
            else if (n == 4) launch(k4);
            else if (n == 8) launch(k8);
            else launch(kn);
        };
        if (numax == 20)
            pick(crt_mfma_kernel<20, 3, 2>, crt_mfma_kernel<20, 3, 4>, crt_mfma_kernel<20, 3, 8>,
                 crt_mfma_kernel<20, 3, 0>);
        else if (numax == 24)
            pick(crt_mfma_kernel<24, 3, 2>, crt_mfma_kernel<24, 3, 4>, crt_mfma_kernel<24, 3, 8>,
                 crt_mfma_kernel<24, 3, 0>);
        else if (numax == 32)
            pick(crt_mfma_kernel<32, 4, 2>, crt_mfma_kernel<32, 4, 4>, crt_mfma_kernel<32, 4, 8>,
                 crt_mfma_kernel<32, 4, 0>);
        else
            pick(crt_mfma_kernel<40, 5, 2>, crt_mfma_kernel<40, 5, 4>, crt_mfma_kernel<40, 5, 8>,
                 crt_mfma_kernel<40, 5, 0>);
    } else if (J <= 32)
        hipLaunchKernelGGL((crt_kernel<32, 256>), dim3(TILE * TILE / 256, ntiles), dim3(256), 0, s,
                           X, L, d_tiles, tiles_m, tiles_n, d_c, d_m, d_qinv, d_q, J, scan_a, scan_b,
                           m, nn, up, pk, n, outw);
    else if (J <= 64)
        hipLaunchKernelGGL((crt_kernel<64, 256>), dim3(TILE * TILE / 256, ntiles), dim3(256), 0, s,
                           X, L, d_tiles, tiles_m, tiles_n, d_c, d_m, d_qinv, d_q, J, scan_a, scan_b,
                           m, nn, up, pk, n, outw);
    else
        hipLaunchKernelGGL((crt_kernel<128, 64>), dim3(TILE * TILE / 64, ntiles), dim3(64), 0, s,
                           X, L, d_tiles, tiles_m, tiles_n, d_c, d_m, d_qinv, d_q, J, scan_a, scan_b,
                           m, nn, up, pk, n, outw);
    HIP_CHECK(hipGetLastError());
    sev.mark(4, s);
    sev.collect(t_stats);  // waits for the stage events when profiling; otherwise nothing waits here
    if (g_profile) {
        // algorithmic work: one exact dot of k terms per output; one INT8
        // GEMM per single-digit modulus and four per two-digit modulus
        const double outs = packed || g.upper ? (double)m * (m + 1) / 2.0 : (double)m * nn;
        const double gemms = (double)single + 4.0 * (L - single);
        t_totals.mp_macs += outs * (double)k;
        t_totals.int8_ops += 2.0 * outs * (double)k * gemms;
        for (int i = 0; i < 4; i++) t_totals.stage_ms[i] += t_stats.stage_ms[i];
        t_totals.calls++;
    }
    return SDPX_GPU_OK;
}

namespace {
// grow-only pinned words for the batched scan summaries (16 an item)
int64_t *batch_words(size_t count) {
    static int64_t *words = nullptr;
    static size_t have = 0;
    if (count > have) {
        if (words) HIP_CHECK(hipHostFree(words));
        words = nullptr;
        have = 0;
        const size_t want = std::max<size_t>(count, 1024);
        HIP_CHECK(hipHostMalloc((void **)&words, want * sizeof(int64_t), hipHostMallocDefault));
        have = want;
    }
    return words;
}
}  // namespace

int rns_gemm_batch(const std::vector<RnsGemmArgs> &items, const std::function<void(size_t)> &after,
                   size_t *failed) {
    Device &dev = device();
    hipStream_t s = dev.stream;
    const size_t N = items.size();
    if (failed) *failed = 0;
    // every item's scans, each into its own rows and summary words
    std::vector<size_t> off_a(N), off_b(N);
    std::vector<char> scanned(N, 0), same(N, 0);
    size_t rows = 0;
    for (size_t i = 0; i < N; i++) {
        const RnsGemmArgs &g = items[i];
        const bool tan = (g.ta == 'N' || g.ta == 'n'), tbn = (g.tb == 'N' || g.tb == 'n');
        const Operand oa{(const uint32_t *)g.a, g.m, g.k, tan ? 1 : g.lda, tan ? g.lda : 1};
        const Operand ob{(const uint32_t *)g.b, g.n, g.k, tbn ? g.ldb : 1, tbn ? 1 : g.ldb};
        same[i] = g.a == g.b && oa.sv == ob.sv && oa.sp == ob.sp && g.m == g.n && g.lda == g.ldb;
        scanned[i] = g.m > 0 && g.n > 0 && g.k > 0;
        if (!scanned[i]) continue;
        off_a[i] = rows;
        rows += (size_t)g.m;
        off_b[i] = same[i] ? off_a[i] : rows;
        if (!same[i]) rows += (size_t)g.n;
    }
    int4 *scan = (int4 *)dev.ws.get(S_BATCH_SCAN, sizeof(int4) * std::max<size_t>(rows, 1));
    int64_t *words = batch_words(16 * std::max<size_t>(N, 1));
    for (size_t i = 0; i < N; i++) {
        if (!scanned[i]) continue;
        const RnsGemmArgs &g = items[i];
        const bool tan = (g.ta == 'N' || g.ta == 'n'), tbn = (g.tb == 'N' || g.tb == 'n');
        const Operand oa{(const uint32_t *)g.a, g.m, g.k, tan ? 1 : g.lda, tan ? g.lda : 1};
        const Operand ob{(const uint32_t *)g.b, g.n, g.k, tbn ? g.ldb : 1, tbn ? 1 : g.ldb};
        launch_scan(oa, g.limbs, scan + off_a[i], s);
        hipLaunchKernelGGL(scan_summary, dim3(1), dim3(256), 0, s, (const int4 *)(scan + off_a[i]), g.m,
                           words + 16 * i);
        if (!same[i]) {
            launch_scan(ob, g.limbs, scan + off_b[i], s);
            hipLaunchKernelGGL(scan_summary, dim3(1), dim3(256), 0, s, (const int4 *)(scan + off_b[i]), g.n,
                               words + 16 * i + 8);
        }
        HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipStreamSynchronize(s));  // the one wait for every window
    for (size_t i = 0; i < N; i++) {
        RnsGemmArgs g = items[i];
        if (scanned[i]) {
            g.scan_a_in = scan + off_a[i];
            g.scan_b_in = scan + off_b[i];
            g.summary = words + 16 * i;
        }
        const int r = rns_gemm_device(g);
        if (r != SDPX_GPU_OK) {
            if (failed) *failed = i;
            return r;
        }
        if (after) after(i);
    }
    return SDPX_GPU_OK;
}

}  // namespace sdpx
