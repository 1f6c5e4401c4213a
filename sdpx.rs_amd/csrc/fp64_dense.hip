// Float64 dense kernels of the Schur path on gfx950: FP64 MFMA GEMM/SYRK
// (v_mfma_f64_16x16x4_f64), blocked Cholesky, triangular solves, and the
// PSD-block Schur assembly S = M^T M with M_b = svec(Rinv A_b Rinv^T).
//
// Reference path: crates/solver/src/solver/kkt/ldl/dense_block.rs
// (tiled_potrf 515-594, tiled_trsv 604-669, solve 795-851: LAPACK/BLAS on
// CPU) and crates/solver/src/solver/kkt/condensed/psd.rs:1432-1456
// (svec(Ginv A Ginv) with Ginv = Rinv^T Rinv, kkt.rs:1119-1127, dotted with
// svec(A_a)).  Float64 results are pinned by tolerance (1e-12 relative),
// not bits: the reference's own BLAS fixes no summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.h"
#include "fp64_dense.h"

namespace sdpx {

typedef double v4d __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// GEMM: C = alpha op(A) op(B) + beta C, batched, optional triangle.
// 64x64 output tile per workgroup, 4 waves of 32x32 (2x2 MFMA 16x16 blocks),
// k staged 16 at a time through LDS with a register prefetch.
// tri: 0 all, 1 only i <= j (upper), 2 only i >= j (lower).
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 16, GPAD = 68;

struct GemmDesc {
    const double *A, *B;
    double *C;
    int64_t m, n, k, lda, ldb, ldc;
    int64_t sa, sb, sc;  // batch strides (elements)
    double alpha, beta;
    int ta, tb;          // 1 when transposed
    int tri;
    int64_t k_split;     // inner elements per split (split-K partials)
    double *partial;     // when splits > 1: [split][m*n] partial products
};

__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmDesc d, int tiles_m, int tiles_n,
                                                       int splits) {
    __shared__ double sA[GK][GPAD];
    __shared__ double sB[GK][GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int t = blockIdx.x;
    const int tm = t % tiles_m;
    t /= tiles_m;
    const int tn = t % tiles_n;
    const int split = t / tiles_n;
    const int batch = blockIdx.y;
    const int64_t i0 = (int64_t)tm * GT, j0 = (int64_t)tn * GT;
    if (d.tri == 1 && i0 > j0 + GT - 1) return;
    if (d.tri == 2 && j0 > i0 + GT - 1) return;
    const double *A = d.A + batch * d.sa;
    const double *B = d.B + batch * d.sb;
    const int64_t kb = (int64_t)split * d.k_split;
    const int64_t ke = min(d.k, kb + d.k_split);

    double ra[4], rb[4];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int e = tid + 256 * u;
            int ii, pp;
            if (!d.ta) { ii = e & 63; pp = e >> 6; } else { pp = e & 15; ii = e >> 4; }
            int64_t gi = i0 + ii, gp = k0 + pp;
            double v = 0.0;
            if (gi < d.m && gp < ke) v = d.ta ? A[gp + gi * d.lda] : A[gi + gp * d.lda];
            ra[u] = v;
            int jj;
            if (!d.tb) { pp = e & 15; jj = e >> 4; } else { jj = e & 63; pp = e >> 6; }
            int64_t gj = j0 + jj;
            gp = k0 + pp;
            v = 0.0;
            if (gj < d.n && gp < ke) v = d.tb ? B[gj + gp * d.ldb] : B[gp + gj * d.ldb];
            rb[u] = v;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int e = tid + 256 * u;
            int ii, pp, jj, qq;
            if (!d.ta) { ii = e & 63; pp = e >> 6; } else { pp = e & 15; ii = e >> 4; }
            sA[pp][ii] = ra[u];
            if (!d.tb) { qq = e & 15; jj = e >> 4; } else { jj = e & 63; qq = e >> 6; }
            sB[qq][jj] = rb[u];
        }
    };
    const int wm = wave >> 1, wn = wave & 1;
    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = v4d{0, 0, 0, 0};
    const int frow = lane & 15, fk = lane >> 4;
    load(kb);
    for (int64_t k0 = kb; k0 < ke; k0 += GK) {
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + GK < ke) load(k0 + GK);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            double fa[2], fb[2];
#pragma unroll
            for (int a = 0; a < 2; a++) fa[a] = sA[kk + fk][wm * 32 + a * 16 + frow];
#pragma unroll
            for (int b = 0; b < 2; b++) fb[b] = sB[kk + fk][wn * 32 + b * 16 + frow];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
    }
    // C/D layout of v_mfma_f64_16x16x4f64: register r of lane l holds row
    // l/16 + 4 r, column l % 16 (one row per register, unlike the
    // four-consecutive-rows grouping of the f32 forms)
    if (splits > 1) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t gi = i0 + wm * 32 + a * 16 + (lane >> 4) + 4 * r;
                    const int64_t gj = j0 + wn * 32 + b * 16 + (lane & 15);
                    if (gi >= d.m || gj >= d.n) continue;
                    if (d.tri == 1 && gi > gj) continue;
                    if (d.tri == 2 && gi < gj) continue;
                    d.partial[(size_t)split * d.m * d.n + gi + gj * d.m] = acc[a][b][r];
                }
        return;
    }
    // every C entry of the tile loaded at once (addresses clamped into the
    // matrix; entries outside it or the triangle are loaded but not
    // stored): a load inside each entry's bounds test would wait for the
    // loads one at a time
    double *Cb = d.C + batch * d.sc;
    double cv[2][2][4];
    if (d.beta != 0.0) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t gi = min(i0 + wm * 32 + a * 16 + (lane >> 4) + 4 * r, d.m - 1);
                    const int64_t gj = min(j0 + wn * 32 + b * 16 + (lane & 15), d.n - 1);
                    cv[a][b][r] = Cb[gi + gj * d.ldc];
                }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t gi = i0 + wm * 32 + a * 16 + (lane >> 4) + 4 * r;
                const int64_t gj = j0 + wn * 32 + b * 16 + (lane & 15);
                const bool keep = gi < d.m && gj < d.n && !(d.tri == 1 && gi > gj) && !(d.tri == 2 && gi < gj);
                const double v = acc[a][b][r];
                const double out = d.beta == 0.0 ? d.alpha * v : fma(d.alpha, v, d.beta * cv[a][b][r]);
                if (keep) Cb[gi + gj * d.ldc] = out;
            }
}

// Deterministic split-K reduction: C = alpha sum_s partial[s] + beta C.
__global__ void splitk_reduce(GemmDesc d, int splits) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= d.m * d.n) return;
    int64_t i = idx % d.m, j = idx / d.m;
    if (d.tri == 1 && i > j) return;
    if (d.tri == 2 && i < j) return;
    double s = 0.0;
    for (int t = 0; t < splits; t++) s += d.partial[(size_t)t * d.m * d.n + idx];
    double *c = d.C + i + j * d.ldc;
    *c = d.beta == 0.0 ? d.alpha * s : fma(d.alpha, s, d.beta * *c);
}

void gemm_f64(const Fp64Gemm &g, hipStream_t s) {
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return;
    GemmDesc d{g.A, g.B, g.C, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.sa, g.sb, g.sc,
               g.alpha, g.beta, g.ta ? 1 : 0, g.tb ? 1 : 0, g.tri, 0, nullptr};
    const int tiles_m = (int)((g.m + GT - 1) / GT), tiles_n = (int)((g.n + GT - 1) / GT);
    // split the inner dimension when the tile grid cannot fill the chip
    int64_t tiles = (int64_t)tiles_m * tiles_n * g.batch;
    int splits = 1;
    if (g.batch == 1 && g.workspace && g.k >= 1024) {
        while (tiles * splits < 512 && g.k / (splits * 2) >= 256 && splits < 32) splits *= 2;
        int64_t need = (int64_t)splits * g.m * g.n;
        if ((size_t)need * 8 > g.workspace_bytes) splits = 1;
    }
    if (g.k == 0) splits = 1;
    d.k_split = splits > 1 ? ((g.k + splits - 1) / splits + GK - 1) / GK * GK : std::max<int64_t>(g.k, 1);
    d.partial = splits > 1 ? g.workspace : nullptr;
    dim3 grid((unsigned)(tiles_m * tiles_n * splits), (unsigned)g.batch);
    hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, s, d, tiles_m, tiles_n, splits);
    HIP_CHECK(hipGetLastError());
    if (splits > 1) {
        int64_t total = g.m * g.n;
        hipLaunchKernelGGL(splitk_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d,
                           splits);
        HIP_CHECK(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------
// Cholesky A = L L^T (lower, column-major), blocked right-looking.
// ---------------------------------------------------------------------------
constexpr int NB = 64;

// Unblocked right-looking Cholesky of one nb x nb diagonal block (nb <= 64).
// Thread (i, g) = (tid % 64, tid / 64) keeps row i's entries of the columns
// c = g + 4 t (t < 16) in registers.  Step j: the threads holding column j
// publish it to LDS (double-buffered by the step's parity), one barrier,
// then every thread forms the pivot's square root and its reciprocal
// itself (the same value everywhere, so the pivot decision needs no second
// barrier), scales its own row's entry and the column entries it needs by
// the reciprocal (LAPACK's dpotf2 scales by 1 / sqrt(d) as well) and
// subtracts the rank-1 term from its entries of the lower triangle.  On a
// non-positive pivot writes info = global column + 1 and stores nothing.
// With the dynamic pivot rule (dense_block.rs:329-365): a pivot d <= eps
// becomes delta unless d < -64 eps_f64 nb |diag_k| (clearly indefinite).
__global__ __launch_bounds__(256) void potrf_diag(double *A, int64_t lda, int64_t j0, int nb,
                                                  int *info, int rule, double eps, double delta,
                                                  const double *diag) {
    __shared__ double col[2][NB];
    const int tid = threadIdx.x, i = tid & 63, g = tid >> 6;
    double x[NB / 4];
#pragma unroll
    for (int t = 0; t < NB / 4; t++) {
        const int c = g + 4 * t;
        x[t] = i < nb && c <= i ? A[(j0 + i) + (j0 + c) * lda] : 0.0;
    }
    if (*info != 0) return;  // the whole workgroup
    int bad = 0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < nb && !bad) {  // uniform: every thread takes the same pivot decisions
            double *cj = col[j & 1];
            if (g == (j & 3)) cj[i] = x[j >> 2];
            __syncthreads();
            const double d = cj[j];
            double djj = 0.0;
            if (!std::isfinite(d)) {
                bad = (int)(j0 + j + 1);
            } else if (rule && d <= eps) {
                if (diag && d < -64.0 * 2.220446049250313e-16 * (double)nb * fabs(diag[j0 + j]))
                    bad = (int)(j0 + j + 1);
                else
                    djj = sqrt(delta);
            } else if (!(d > 0.0)) {
                bad = (int)(j0 + j + 1);
            } else {
                djj = sqrt(d);
            }
            if (!bad) {
                const double r = 1.0 / djj;
                const double lij = cj[i] * r;
                // every column entry read unconditionally and the update
                // selected per entry: a branch around each LDS read would
                // wait for the reads one at a time (entries past nb are
                // never selected, so their contents do not matter)
                // (columns c = g + 4 t <= 4 t + 3 <= j are finished: skipped
                // at compile time, the loops being unrolled over j and t)
                double lc[NB / 4];
#pragma unroll
                for (int t = 0; t < NB / 4; t++) lc[t] = cj[g + 4 * t];
#pragma unroll
                for (int t = 0; t < NB / 4; t++) {
                    const int c = g + 4 * t;
                    const double u = fma(-lij, lc[t] * r, x[t]);
                    x[t] = c > j && c <= i ? u : x[t];  // c <= i implies i > j
                }
                if (g == (j & 3) && i >= j) x[j >> 2] = i == j ? djj : lij;
            }
        }
    }
    if (bad) {
        if (tid == 0) *info = bad;
        return;
    }
#pragma unroll
    for (int t = 0; t < NB / 4; t++) {
        const int c = g + 4 * t;
        if (i < nb && c <= i) A[(j0 + i) + (j0 + c) * lda] = x[t];
    }
}

__device__ __forceinline__ double readlane_f64(double v, int k) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)b, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// X L^T = B for the rows below a diagonal block: row panel rows r0..r1,
// columns j0..j0+nb.  One wave a row (the grid's waves stride over the
// rows), lane j holding x_j: step j takes x_j from lane j (v_readlane)
// times 1/L_jj (reference BLAS dtrsm scales by the reciprocal as well) and
// every later lane subtracts its multiple, so each x_j sees the updates of
// a row-by-row substitution in the same order.
constexpr int TRSM_PANEL_GRID = 1024;
__global__ __launch_bounds__(256) void trsm_panel(double *A, int64_t lda, int64_t j0, int nb,
                                                  int64_t r0, int64_t r1, const int *info) {
    __shared__ double L[NB][NB + 1];
    __shared__ double rinv[NB];
    if (*info != 0) return;
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += 256) {
        int i = e % NB, j = e / NB;
        L[i][j] = i < nb && j < nb && i >= j ? A[(j0 + i) + (j0 + j) * lda] : 0.0;
    }
    __syncthreads();
    if (tid < nb) rinv[tid] = 1.0 / L[tid][tid];
    __syncthreads();
    const int lane = tid & 63;
    for (int64_t row = r0 + (int64_t)blockIdx.x * 4 + (tid >> 6); row < r1; row += (int64_t)gridDim.x * 4) {
        double x = lane < nb ? A[row + (j0 + lane) * lda] : 0.0;
        for (int j = 0; j < nb; j++) {
            const double xj = readlane_f64(x, j) * rinv[j];
            const double u = fma(-xj, L[lane][j], x);  // read for every lane: no branch around the LDS read
            x = lane == j ? xj : (lane > j ? u : x);
        }
        if (lane < nb) A[row + (j0 + lane) * lda] = x;
    }
}

// A second stream (default priority) and two events per device for the
// factor's look-ahead, created on first use and kept until the process ends.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t panel = nullptr, update = nullptr;
};
static SideStream &side_stream() {
    static SideStream side[64];
    static std::mutex mu;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    SideStream &x = side[dev & 63];
    if (!x.s) {
        HIP_CHECK(hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&x.panel, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&x.update, hipEventDisableTiming));
    }
    return x;
}

static void potrf_panel(double *A, int64_t n, int64_t lda, int64_t j0, int *d_info, hipStream_t s,
                        const PivotRule *rule) {
    const int nb = (int)std::min<int64_t>(NB, n - j0);
    hipLaunchKernelGGL(potrf_diag, dim3(1), dim3(256), 0, s, A, lda, j0, nb, d_info, rule ? 1 : 0,
                       rule ? rule->eps : 0.0, rule ? rule->delta : 0.0, rule ? rule->diag : nullptr);
    HIP_CHECK(hipGetLastError());
    const int64_t r0 = j0 + nb;
    if (r0 >= n) return;
    hipLaunchKernelGGL(trsm_panel, dim3((unsigned)std::min<int64_t>((n - r0 + 3) / 4, TRSM_PANEL_GRID)),
                       dim3(256), 0, s, A, lda, j0, nb, r0, n, d_info);
    HIP_CHECK(hipGetLastError());
}

// A[r0:, c0:c0+nc] -= A[r0:, j0:j0+NB] A[c0:c0+nc, j0:j0+NB]^T over the
// lower triangle (r0 == c0: the columns' diagonal block and everything below)
static void panel_update(double *A, int64_t n, int64_t lda, int64_t j0, int64_t c0, int64_t nc, double *ws,
                         size_t ws_bytes, hipStream_t s) {
    Fp64Gemm g;
    g.m = n - c0;
    g.n = nc;
    g.k = NB;
    g.A = A + c0 + j0 * lda;
    g.lda = lda;
    g.B = A + c0 + j0 * lda;
    g.ldb = lda;
    g.tb = true;
    g.C = A + c0 + c0 * lda;
    g.ldc = lda;
    g.alpha = -1.0;
    g.beta = 1.0;
    g.tri = 2;
    g.workspace = ws;
    g.workspace_bytes = ws_bytes;
    gemm_f64(g, s);
}

// Blocked right-looking Cholesky, with look-ahead from order 4096: once
// panel p is factored (diagonal block, then the rows below it), the next
// panel's columns take panel p's update on the caller's stream and are
// factored there at once, while the rest of the trailing matrix takes the
// same update on a second
// stream; the next panel's update waits for that.  So the diagonal factor
// and the panel solve of each step (latency-bound single-block kernels) run
// beside the step's large update instead of between updates.  Every entry
// receives the same updates in the same order as without look-ahead.  The
// two cross-stream waits a step cost more than they hide below order 4096
// (measured: order 8192 18.4 -> 16.2 ms, order 2000 2.18 -> 2.41 ms).
constexpr int64_t LOOKAHEAD_MIN_ORDER = 4096;
void potrf_f64(double *A, int64_t n, int64_t lda, int *d_info, double *ws, size_t ws_bytes,
               hipStream_t s, const PivotRule *rule) {
    if (n < LOOKAHEAD_MIN_ORDER) {
        for (int64_t j0 = 0; j0 < n; j0 += NB) {
            potrf_panel(A, n, lda, j0, d_info, s, rule);
            const int64_t r0 = j0 + NB;
            if (r0 < n) panel_update(A, n, lda, j0, r0, n - r0, ws, ws_bytes, s);
        }
        return;
    }
    SideStream &side = side_stream();
    potrf_panel(A, n, lda, 0, d_info, s, rule);
    for (int64_t j0 = 0; j0 + NB < n; j0 += NB) {
        const int64_t r0 = j0 + NB, nb1 = std::min<int64_t>(NB, n - r0), r1 = r0 + nb1;
        if (r1 < n) {  // the rest of the trailing matrix on the second stream (split-K off: no workspace)
            HIP_CHECK(hipEventRecord(side.panel, s));
            HIP_CHECK(hipStreamWaitEvent(side.s, side.panel, 0));
            panel_update(A, n, lda, j0, r1, n - r1, nullptr, 0, side.s);
            HIP_CHECK(hipEventRecord(side.update, side.s));
        }
        panel_update(A, n, lda, j0, r0, nb1, ws, ws_bytes, s);
        potrf_panel(A, n, lda, r0, d_info, s, rule);
        if (r1 < n) HIP_CHECK(hipStreamWaitEvent(s, side.update, 0));
    }
}

// Forward/back substitution with the factor for nrhs columns:
// lower: solve L Y = B; upper (trans): solve L^T X = Y, one 64-row diagonal
// block a launch (then a GEMM update of the remaining rows).  One wave a
// right-hand side, lane i holding row i: step i takes row i from lane i
// (v_readlane) times the diagonal's reciprocal and every later (earlier,
// for the transpose) row subtracts its multiple, so each row sees the same
// sequence of updates as a row-by-row substitution; four columns a
// 256-thread workgroup.
__global__ __launch_bounds__(256) void trsv_block(const double *L, int64_t ldl, double *B,
                                                  int64_t ldb, int64_t j0, int nb, int64_t nrhs,
                                                  int trans) {
    __shared__ double s[NB][NB + 1];
    __shared__ double rinv[NB];
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += 256) {
        int i = e % nb, j = e / nb;
        s[i][j] = i >= j ? L[(j0 + i) + (j0 + j) * ldl] : 0.0;
    }
    __syncthreads();
    if (tid < nb) rinv[tid] = 1.0 / s[tid][tid];
    __syncthreads();
    const int lane = tid & 63;
    const int64_t c = (int64_t)blockIdx.x * 4 + (tid >> 6);
    if (c >= nrhs) return;  // whole waves
    double x = lane < nb ? B[(j0 + lane) + c * ldb] : 0.0;
    // the LDS entry is read for every lane and the update selected, so no
    // branch waits around the read (rows past nb are never selected)
    if (!trans) {
        for (int i = 0; i < nb; i++) {
            const double xi = readlane_f64(x, i) * rinv[i];
            const double u = fma(-s[lane][i], xi, x);
            x = lane == i ? xi : (lane > i && lane < nb ? u : x);
        }
    } else {
        for (int ii = nb - 1; ii >= 0; ii--) {
            const double xi = readlane_f64(x, ii) * rinv[ii];
            const double u = fma(-s[ii][lane], xi, x);
            x = lane == ii ? xi : (lane < ii ? u : x);
        }
    }
    if (lane < nb) B[(j0 + lane) + c * ldb] = x;
}

void trsm_lower_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs, int64_t ldb,
                    hipStream_t s) {
    const unsigned cgrid = (unsigned)((nrhs + 3) / 4);
    for (int64_t j0 = 0; j0 < n; j0 += NB) {
        const int nb = (int)std::min<int64_t>(NB, n - j0);
        hipLaunchKernelGGL(trsv_block, dim3(cgrid), dim3(256), 0, s, L, ldl, B, ldb, j0, nb, nrhs, 0);
        HIP_CHECK(hipGetLastError());
        if (j0 + nb < n) {
            Fp64Gemm g;  // B[j0+nb:] -= L[j0+nb:, j0:j0+nb] B[j0:j0+nb]
            g.m = n - j0 - nb;
            g.n = nrhs;
            g.k = nb;
            g.A = L + (j0 + nb) + j0 * ldl;
            g.lda = ldl;
            g.B = B + j0;
            g.ldb = ldb;
            g.C = B + j0 + nb;
            g.ldc = ldb;
            g.alpha = -1.0;
            g.beta = 1.0;
            gemm_f64(g, s);
        }
    }
}

void trsm_lower_t_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs,
                      int64_t ldb, hipStream_t s) {
    const unsigned cgrid = (unsigned)((nrhs + 3) / 4);
    int64_t last = ((n - 1) / NB) * NB;
    for (int64_t j0 = last; j0 >= 0; j0 -= NB) {
        const int nb = (int)std::min<int64_t>(NB, n - j0);
        hipLaunchKernelGGL(trsv_block, dim3(cgrid), dim3(256), 0, s, L, ldl, B, ldb, j0, nb, nrhs, 1);
        HIP_CHECK(hipGetLastError());
        if (j0 > 0) {
            Fp64Gemm g;  // B[0:j0] -= L[j0:j0+nb, 0:j0]^T B[j0:j0+nb]
            g.m = j0;
            g.n = nrhs;
            g.k = nb;
            g.A = L + j0;
            g.lda = ldl;
            g.ta = true;
            g.B = B + j0;
            g.ldb = ldb;
            g.C = B;
            g.ldc = ldb;
            g.alpha = -1.0;
            g.beta = 1.0;
            gemm_f64(g, s);
        }
    }
}

void potrs_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs, int64_t ldb,
               double *ws, size_t ws_bytes, hipStream_t s) {
    (void)ws;
    (void)ws_bytes;
    if (n == 0 || nrhs == 0) return;
    trsm_lower_f64(L, n, ldl, B, nrhs, ldb, s);
    trsm_lower_t_f64(L, n, ldl, B, nrhs, ldb, s);
}

// ---------------------------------------------------------------------------
// PSD-block Schur assembly.
// ---------------------------------------------------------------------------
// svec columns (upper-triangular column order, off-diagonals scaled by
// sqrt 2) -> dense symmetric matrices stacked vertically: Astack[b*n + i, j].
__global__ void svec_unpack(const double *svec, int64_t s, int64_t n, int64_t m, double *out) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= s * m) return;
    int64_t b = idx / s, e = idx % s;
    // e -> (i, j), i <= j, column-major upper
    int64_t j = (int64_t)((sqrt(8.0 * (double)e + 1.0) - 1.0) / 2.0);
    while (j * (j + 1) / 2 > e) j--;
    while ((j + 1) * (j + 2) / 2 <= e) j++;
    int64_t i = e - j * (j + 1) / 2;
    double v = svec[e + b * s];
    if (i != j) v *= 0.70710678118654752440;
    const int64_t ld = n * m;
    out[(b * n + i) + j * ld] = v;
    out[(b * n + j) + i * ld] = v;
}

// U_b (stacked, ld = n*m) -> svec columns M[:, b] (sqrt 2 off-diagonal),
// symmetrised as (U_ij + U_ji)/2.
__global__ void svec_pack(const double *U, int64_t s, int64_t n, int64_t m, double *M) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= s * m) return;
    int64_t b = idx / s, e = idx % s;
    int64_t j = (int64_t)((sqrt(8.0 * (double)e + 1.0) - 1.0) / 2.0);
    while (j * (j + 1) / 2 > e) j--;
    while ((j + 1) * (j + 2) / 2 <= e) j++;
    int64_t i = e - j * (j + 1) / 2;
    const int64_t ld = n * m;
    double v = i == j ? U[(b * n + i) + j * ld]
                      : (U[(b * n + i) + j * ld] + U[(b * n + j) + i * ld]) * 0.70710678118654752440;
    M[e + b * s] = v;
}

}  // namespace sdpx

namespace sdpx {

// Workspace slots 30..35 are this file's.
void schur_psd_f64(int64_t n, int64_t m, const double *rinv, const double *a_svec, double *S,
                   int64_t lds, hipStream_t st) {
    Workspace &ws = device().ws;
    const int64_t s = n * (n + 1) / 2;
    double *stack = (double *)ws.get(30, sizeof(double) * n * n * m);
    double *x = (double *)ws.get(31, sizeof(double) * n * n * m);
    double *M = (double *)ws.get(32, sizeof(double) * s * m);
    const int64_t total = s * m;
    hipLaunchKernelGGL(svec_unpack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a_svec,
                       s, n, m, stack);
    HIP_CHECK(hipGetLastError());
    // X = Astack Rinv^T  (nm x n): X_b = A_b Rinv^T
    Fp64Gemm g1;
    g1.m = n * m;
    g1.n = n;
    g1.k = n;
    g1.A = stack;
    g1.lda = n * m;
    g1.B = rinv;
    g1.ldb = n;
    g1.tb = true;
    g1.C = x;
    g1.ldc = n * m;
    gemm_f64(g1, st);
    // U_b = Rinv X_b (batched; written over the stack buffer)
    Fp64Gemm g2;
    g2.m = n;
    g2.n = n;
    g2.k = n;
    g2.A = rinv;
    g2.lda = n;
    g2.B = x;
    g2.ldb = n * m;
    g2.sb = n;
    g2.C = stack;
    g2.ldc = n * m;
    g2.sc = n;
    g2.batch = m;
    gemm_f64(g2, st);
    hipLaunchKernelGGL(svec_pack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, stack, s,
                       n, m, M);
    HIP_CHECK(hipGetLastError());
    // S = M^T M (upper)
    size_t wsb = sizeof(double) * (size_t)m * m * 16;
    double *part = (double *)ws.get(33, wsb);
    Fp64Gemm g3;
    g3.m = m;
    g3.n = m;
    g3.k = s;
    g3.A = M;
    g3.lda = s;
    g3.ta = true;
    g3.B = M;
    g3.ldb = s;
    g3.C = S;
    g3.ldc = lds;
    g3.tri = 1;
    g3.workspace = part;
    g3.workspace_bytes = wsb;
    gemm_f64(g3, st);
}

}  // namespace sdpx

namespace sdpx {

// Lower triangle <- upper triangle (the factor reads the upper triangle, as
// the reference's dense factor takes the KKT's upper CSC).
__global__ void mirror_upper_to_lower(double *A, int64_t n, int64_t lda) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    int64_t i = idx % n, j = idx / n;
    if (i > j) A[i + j * lda] = A[j + i * lda];
}

void mirror_upper(double *A, int64_t n, int64_t lda, hipStream_t s) {
    hipLaunchKernelGGL(mirror_upper_to_lower, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, s,
                       A, n, lda);
    HIP_CHECK(hipGetLastError());
}

}  // namespace sdpx

namespace sdpx {

// Adjoint of the Schur solve at a PSD block (no reference counterpart; the
// reference has no differentiation).  At the IPM solution the condensed
// Newton system is M dy = db + A(W dC W) with M_ij = <A_i, W A_j W>,
// W = Rinv^T Rinv.  For cotangents ybar of y*: v = M^-T ybar = M^-1 ybar
// (the same factor: M is symmetric), bbar = v, Cbar = W A*(v) W.
// ybar/bgrad: m x nrhs; cgrad: s x nrhs svec columns (device pointers).
void adjoint_psd_f64(const double *Lm, int64_t m, int64_t n, const double *rinv,
                     const double *a_svec, int64_t nrhs, const double *ybar, double *bgrad,
                     double *cgrad, hipStream_t st) {
    Workspace &ws = device().ws;
    const int64_t s = n * (n + 1) / 2;
    // v = M^-1 ybar (into bgrad)
    HIP_CHECK(hipMemcpyAsync(bgrad, ybar, 8 * (size_t)m * nrhs, hipMemcpyDeviceToDevice, st));
    potrs_f64(Lm, m, m, bgrad, nrhs, m, nullptr, 0, st);
    // Z = A_svec v  (s x nrhs)
    double *z = (double *)ws.get(60, 8 * (size_t)s * nrhs);
    Fp64Gemm g0;
    g0.m = s; g0.n = nrhs; g0.k = m;
    g0.A = a_svec; g0.lda = s; g0.B = bgrad; g0.ldb = m; g0.C = z; g0.ldc = s;
    gemm_f64(g0, st);
    double *stack = (double *)ws.get(61, 8 * (size_t)n * n * nrhs);
    double *x = (double *)ws.get(62, 8 * (size_t)n * n * nrhs);
    const int64_t total = s * nrhs;
    hipLaunchKernelGGL(svec_unpack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, z, s, n,
                       nrhs, stack);
    HIP_CHECK(hipGetLastError());
    // W Z W = Rinv^T (Rinv Z Rinv^T) Rinv, two congruences on the stack
    for (int pass = 0; pass < 2; pass++) {
        const bool t = pass == 1;  // second pass uses Rinv^T
        Fp64Gemm g1;  // X_k = Z_k F^T  (F = Rinv, then Rinv^T)
        g1.m = n * nrhs; g1.n = n; g1.k = n;
        g1.A = stack; g1.lda = n * nrhs; g1.B = rinv; g1.ldb = n; g1.tb = !t;
        g1.C = x; g1.ldc = n * nrhs;
        gemm_f64(g1, st);
        Fp64Gemm g2;  // Z_k = F X_k
        g2.m = n; g2.n = n; g2.k = n;
        g2.A = rinv; g2.lda = n; g2.ta = t; g2.B = x; g2.ldb = n * nrhs; g2.sb = n;
        g2.C = stack; g2.ldc = n * nrhs; g2.sc = n; g2.batch = nrhs;
        gemm_f64(g2, st);
    }
    hipLaunchKernelGGL(svec_pack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, stack, s, n,
                       nrhs, cgrad);
    HIP_CHECK(hipGetLastError());
}

}  // namespace sdpx

namespace sdpx {
namespace {
// W X_k W for nrhs symmetric matrices given as svec columns (in: s x nrhs),
// W = Rinv^T Rinv, through two congruences of the stacked matrices
void w_congruence(const double *rinv, int64_t n, const double *in, int64_t nrhs, double *out, double *stack,
                  double *x, hipStream_t st) {
    const int64_t s = n * (n + 1) / 2, total = s * nrhs;
    hipLaunchKernelGGL(svec_unpack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, in, s, n, nrhs, stack);
    HIP_CHECK(hipGetLastError());
    for (int pass = 0; pass < 2; pass++) {
        const bool t = pass == 1;
        Fp64Gemm g1;  // X_k = Z_k F^T (F = Rinv, then Rinv^T)
        g1.m = n * nrhs; g1.n = n; g1.k = n;
        g1.A = stack; g1.lda = n * nrhs; g1.B = rinv; g1.ldb = n; g1.tb = !t;
        g1.C = x; g1.ldc = n * nrhs;
        gemm_f64(g1, st);
        Fp64Gemm g2;  // Z_k = F X_k
        g2.m = n; g2.n = n; g2.k = n;
        g2.A = rinv; g2.lda = n; g2.ta = t; g2.B = x; g2.ldb = n * nrhs; g2.sb = n;
        g2.C = stack; g2.ldc = n * nrhs; g2.sc = n; g2.batch = nrhs;
        gemm_f64(g2, st);
    }
    hipLaunchKernelGGL(svec_pack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, stack, s, n, nrhs, out);
    HIP_CHECK(hipGetLastError());
}

// out[:, k] = a[:, k] + w_k b  (s x nrhs; a may be null: zero)
__global__ void axpy_cols(const double *a, const double *wts, const double *b, int64_t s, int64_t nrhs, double *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= s * nrhs) return;
    const int64_t i = e % s, k = e / s;
    out[e] = (a ? a[e] : 0.0) + (wts ? wts[k] * b[i] : 0.0);
}

// cgrad[:, k] = U[:, k] - Z[:, k] + w_k x
__global__ void cgrad_cols(const double *U, const double *Z, const double *wts, const double *x, int64_t s,
                           int64_t nrhs, double *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= s * nrhs) return;
    const int64_t i = e % s, k = e / s;
    out[e] = U[e] - Z[e] + (wts ? wts[k] * x[i] : 0.0);
}

// agrad[i + s (j + m k)] = y_j (Z - U)[i, k] - v[j, k] x_i
__global__ void agrad_kernel(const double *U, const double *Z, const double *y, const double *v, const double *x,
                             int64_t s, int64_t m, int64_t nrhs, double *agrad) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= s * m * nrhs) return;
    const int64_t i = e % s, j = (e / s) % m, k = e / (s * m);
    agrad[e] = y[j] * (Z[i + s * k] - U[i + s * k]) - v[j + m * k] * x[i];
}
}  // namespace

// Full adjoint at a central-path point (no reference counterpart).  The
// point (X, y, S) with A(X) = b, A*(y) + S = C, X S = mu I moves with the
// data as dX = -W dS W (W = Rinv^T Rinv the NT scaling: W S W = X), dS = dC
// - dA*(y) - A*(dy), and M dy = db + A(W (dC - dA*(y)) W) - dA(X), M_ij =
// <A_i, W A_j W>.  For the functional L = ybar . y + <Xbar, X> + obj <C, X>,
// with Z = W (Xbar + obj C) W, g = ybar + A(Z), v = M^-1 g (the same
// factor: M is symmetric) and U = W A*(v) W:
//   dL/db = v,  dL/dC = U - Z + obj X,  dL/dA_i = y_i (Z - U) - v_i X.
// All operands device pointers; matrices as svec columns (s = n(n+1)/2).
void adjoint_psd_full_f64(const double *Lm, int64_t m, int64_t n, const double *rinv, const double *a_svec,
                          const double *y, const double *x_svec, const double *c_svec, int64_t nrhs,
                          const double *ybar, const double *xbar, const double *obj, double *bgrad,
                          double *cgrad, double *agrad, hipStream_t st) {
    Workspace &ws = device().ws;
    const int64_t s = n * (n + 1) / 2, total = s * nrhs;
    double *z = (double *)ws.get(60, 8 * (size_t)s * nrhs);
    double *u = (double *)ws.get(68, 8 * (size_t)s * nrhs);
    double *stack = (double *)ws.get(61, 8 * (size_t)n * n * nrhs);
    double *x = (double *)ws.get(62, 8 * (size_t)n * n * nrhs);
    // Z = W (Xbar + obj C) W
    const bool have_z = xbar || obj;
    if (have_z) {
        hipLaunchKernelGGL(axpy_cols, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, xbar, obj, c_svec, s,
                           nrhs, u);
        HIP_CHECK(hipGetLastError());
        w_congruence(rinv, n, u, nrhs, z, stack, x, st);
    } else {
        HIP_CHECK(hipMemsetAsync(z, 0, 8 * (size_t)total, st));
    }
    // g = ybar + A(Z) into bgrad, then v = M^-1 g in place
    if (ybar)
        HIP_CHECK(hipMemcpyAsync(bgrad, ybar, 8 * (size_t)m * nrhs, hipMemcpyDeviceToDevice, st));
    if (have_z || !ybar) {
        Fp64Gemm g0;
        g0.m = m; g0.n = nrhs; g0.k = s;
        g0.A = a_svec; g0.lda = s; g0.ta = true; g0.B = z; g0.ldb = s; g0.C = bgrad; g0.ldc = m;
        g0.beta = ybar ? 1.0 : 0.0;
        gemm_f64(g0, st);
    }
    potrs_f64(Lm, m, m, bgrad, nrhs, m, nullptr, 0, st);
    // U = W A*(v) W
    Fp64Gemm g1;
    g1.m = s; g1.n = nrhs; g1.k = m;
    g1.A = a_svec; g1.lda = s; g1.B = bgrad; g1.ldb = m; g1.C = u; g1.ldc = s;
    gemm_f64(g1, st);
    double *wu = cgrad;  // W A*(v) W, then U - Z + obj X in place
    w_congruence(rinv, n, u, nrhs, wu, stack, x, st);
    HIP_CHECK(hipMemcpyAsync(u, wu, 8 * (size_t)total, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(cgrad_cols, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, u, z, obj, x_svec, s,
                       nrhs, cgrad);
    HIP_CHECK(hipGetLastError());
    if (agrad) {
        const int64_t ta = s * m * nrhs;
        hipLaunchKernelGGL(agrad_kernel, dim3((unsigned)((ta + 255) / 256)), dim3(256), 0, st, u, z, y, bgrad, x_svec,
                           s, m, nrhs, agrad);
        HIP_CHECK(hipGetLastError());
    }
}

}  // namespace sdpx

namespace sdpx {
namespace {
// S[i + j lds] += P (CSC, entries i <= j), one thread a column
__global__ void add_csc_upper(const int64_t *colptr, const int64_t *rowval, const double *nzval, int64_t m,
                              double *S, int64_t lds) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    for (int64_t p = colptr[j]; p < colptr[j + 1]; p++) {
        const int64_t i = rowval[p];
        if (i <= j) S[i + j * lds] += nzval[p];
    }
}

// S[c_a, c_b] += T[a, b] for a <= b (T upper, ld k): a block's columns
// scattered into the reduced Schur's upper triangle
__global__ void scatter_block_upper(const double *T, int64_t k, const uint64_t *cols, double *S, int64_t lds) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k * k) return;
    const int64_t a = e % k, b = e / k;
    if (a > b) return;
    const uint64_t ca = cols[a], cb = cols[b];
    const uint64_t i = ca < cb ? ca : cb, j = ca < cb ? cb : ca;
    S[i + j * lds] += T[a + b * k];
}

// G[r + c ld] = w_r B[r + c ld]
__global__ void scale_rows(const double *B, int64_t ld, const double *w, int64_t r, int64_t m, double *G) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r * m) return;
    const int64_t i = e % r, c = e / r;
    G[i + c * ld] = w[i] * B[i + c * ld];
}

__global__ void zero_upper(double *S, int64_t m, int64_t lds) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * m) return;
    const int64_t i = e % m, j = e / m;
    if (i <= j) S[i + j * lds] = 0.0;
}
}  // namespace

// The condensed Schur assembly of a reduced KKT (device pointers): S's upper
// triangle = P, then each PSD block's contribution S_ab = <A_a, W A_b W>
// (W = Rinv^T Rinv) scattered to its global columns in cone order, then the
// orthant rows' Gram B^T diag(w) B.  Workspace slots 34-36.
// out[i] = x[cols[i]] (k entries)
__global__ void gather_cols(const double *x, const uint64_t *cols, int64_t k, double *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < k) out[e] = x[cols[e]];
}

// x[cols[i]] += u[i] (k entries; the columns of one block are distinct)
__global__ void scatter_add_cols(const double *u, const uint64_t *cols, int64_t k, double *x) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < k) x[cols[e]] += u[e];
}

// v = -v (n entries)
__global__ void negate_vec(double *v, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) v[e] = -v[e];
}

// out = w o v (r entries)
__global__ void scale_vec(const double *w, const double *v, int64_t r, double *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < r) out[e] = w[e] * v[e];
}

// The eliminated rows' inverse scaling H^-1 of one PSD block on an svec
// vector: svec(W smat(v) W), W = Rinv^T Rinv, as two congruences with Rinv
// (the binary64 form of PsdScaling::apply with inverse set,
// kkt/condensed/psd.rs:1968-2000).
static void psd_hinv(const PsdBlockF64 &b, const double *v, double *out, hipStream_t st) {
    Workspace &ws = device().ws;
    double *stack = (double *)ws.get(36, sizeof(double) * b.n * b.n);
    double *x = (double *)ws.get(37, sizeof(double) * b.n * b.n);
    w_congruence(b.rinv, b.n, v, 1, out, stack, x, st);
}

void condensed_rhs_f64(int64_t m, const std::vector<PsdBlockF64> &blocks, int64_t rows, const double *B, int64_t ldb,
                       const double *w, const double *bx, const double *bz, double *xr, hipStream_t st) {
    Workspace &ws = device().ws;
    HIP_CHECK(hipMemcpyAsync(xr, bx, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    int64_t off = 0;
    for (const auto &b : blocks) {
        const int64_t s = b.n * (b.n + 1) / 2;
        if (b.ncols > 0) {
            double *t = (double *)ws.get(38, sizeof(double) * s);
            double *u = (double *)ws.get(39, sizeof(double) * b.ncols);
            psd_hinv(b, bz + off, t, st);
            Fp64Gemm g;  // u = A_b^T t
            g.m = b.ncols; g.n = 1; g.k = s;
            g.A = b.a_svec; g.lda = s; g.ta = true; g.B = t; g.ldb = s;
            g.C = u; g.ldc = b.ncols;
            gemm_f64(g, st);
            hipLaunchKernelGGL(scatter_add_cols, dim3((unsigned)((b.ncols + 255) / 256)), dim3(256), 0, st, u,
                               b.columns, b.ncols, xr);
            HIP_CHECK(hipGetLastError());
        }
        off += s;
    }
    if (rows > 0) {
        double *t = (double *)ws.get(38, sizeof(double) * rows);
        hipLaunchKernelGGL(scale_vec, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, w, bz + off, rows, t);
        HIP_CHECK(hipGetLastError());
        Fp64Gemm g;  // xr += B^T t
        g.m = m; g.n = 1; g.k = rows;
        g.A = B; g.lda = ldb; g.ta = true; g.B = t; g.ldb = rows;
        g.C = xr; g.ldc = m; g.beta = 1.0;
        gemm_f64(g, st);
    }
}

void condensed_recover_f64(int64_t m, const std::vector<PsdBlockF64> &blocks, int64_t rows, const double *B,
                           int64_t ldb, const double *w, const double *x, const double *bz, double *z, hipStream_t st) {
    Workspace &ws = device().ws;
    int64_t off = 0;
    for (const auto &b : blocks) {
        const int64_t s = b.n * (b.n + 1) / 2;
        double *v = (double *)ws.get(38, sizeof(double) * s);
        HIP_CHECK(hipMemcpyAsync(v, bz + off, sizeof(double) * s, hipMemcpyDeviceToDevice, st));
        if (b.ncols > 0) {
            double *g0 = (double *)ws.get(39, sizeof(double) * b.ncols);
            hipLaunchKernelGGL(gather_cols, dim3((unsigned)((b.ncols + 255) / 256)), dim3(256), 0, st, x, b.columns,
                               b.ncols, g0);
            HIP_CHECK(hipGetLastError());
            Fp64Gemm g;  // v = A_b x[cols] - bz_b
            g.m = s; g.n = 1; g.k = b.ncols;
            g.A = b.a_svec; g.lda = s; g.B = g0; g.ldb = b.ncols;
            g.C = v; g.ldc = s; g.beta = -1.0;
            gemm_f64(g, st);
        } else {  // no constraint touches the block: v = -bz_b
            hipLaunchKernelGGL(negate_vec, dim3((unsigned)((s + 255) / 256)), dim3(256), 0, st, v, s);
            HIP_CHECK(hipGetLastError());
        }
        psd_hinv(b, v, z + off, st);
        off += s;
    }
    if (rows > 0) {
        double *v = (double *)ws.get(38, sizeof(double) * rows);
        HIP_CHECK(hipMemcpyAsync(v, bz + off, sizeof(double) * rows, hipMemcpyDeviceToDevice, st));
        Fp64Gemm g;  // v = B x - bz_orth
        g.m = rows; g.n = 1; g.k = m;
        g.A = B; g.lda = ldb; g.B = x; g.ldb = m;
        g.C = v; g.ldc = rows; g.beta = -1.0;
        gemm_f64(g, st);
        hipLaunchKernelGGL(scale_vec, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, w, v, rows, z + off);
        HIP_CHECK(hipGetLastError());
    }
}

void condensed_assemble_f64(int64_t m, const int64_t *p_colptr, const int64_t *p_rowval, const double *p_nzval,
                            const std::vector<PsdBlockF64> &blocks, int64_t rows, const double *B, int64_t ldb,
                            const double *w, double *S, int64_t lds, hipStream_t st) {
    Workspace &ws = device().ws;
    hipLaunchKernelGGL(zero_upper, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, st, S, m, lds);
    HIP_CHECK(hipGetLastError());
    if (p_colptr) {
        hipLaunchKernelGGL(add_csc_upper, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, p_colptr, p_rowval,
                           p_nzval, m, S, lds);
        HIP_CHECK(hipGetLastError());
    }
    for (const auto &b : blocks) {
        if (b.ncols == 0) continue;
        double *T = (double *)ws.get(34, 8 * (size_t)b.ncols * b.ncols);
        schur_psd_f64(b.n, b.ncols, b.rinv, b.a_svec, T, b.ncols, st);
        const int64_t kk = b.ncols * b.ncols;
        hipLaunchKernelGGL(scatter_block_upper, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0, st, T, b.ncols,
                           b.columns, S, lds);
        HIP_CHECK(hipGetLastError());
    }
    if (rows > 0 && B) {
        double *G = (double *)ws.get(35, 8 * (size_t)ldb * m);
        hipLaunchKernelGGL(scale_rows, dim3((unsigned)((rows * m + 255) / 256)), dim3(256), 0, st, B, ldb, w, rows, m,
                           G);
        HIP_CHECK(hipGetLastError());
        Fp64Gemm g;
        g.m = m; g.n = m; g.k = rows;
        g.A = B; g.lda = ldb; g.ta = true; g.B = G; g.ldb = ldb; g.C = S; g.ldc = lds;
        g.beta = 1.0; g.tri = 1;
        gemm_f64(g, st);
    }
}

}  // namespace sdpx
