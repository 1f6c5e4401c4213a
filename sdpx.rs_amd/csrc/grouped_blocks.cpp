// Grouped exact blocks (rns_blas::gemm_blocks_upper, rns_blas.rs:2671-2922,
// behind XgemmScalar::xgemm_blocks_upper_exact, traits.rs:250-258) on the
// device: every block's rows scattered at its global columns into two dense
// K x m operands in HBM (Y, and Z = RN(diag(d) Y) entrywise), zeros
// elsewhere, then one exact product Z^T Y whose upper entries are each the
// exact sum over all blocks rounded once.  Zeros outside a block's columns
// add nothing to any exact sum, so one launch over the stacked rows is the
// grouped product of blocks of uneven height.
#include <algorithm>
#include <functional>
#include <vector>

#include "common.h"
#include "mp_kernels.h"
#include "rns_gemm.h"

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);

namespace {
enum { GB_Z = 150, GB_Y, GB_VALS, GB_COLS, GB_DESC, GB_OUT };
}

// struct sizes, arrays present, columns ascending below m; K = total rows
int32_t check_exact_blocks(uint64_t m, uint64_t nblocks, const sdpx_gpu_exact_block *blocks, uint64_t &K) {
    if (nblocks && !blocks) return abi_invalid("null blocks");
    K = 0;
    for (uint64_t b = 0; b < nblocks; b++) {
        const auto &bl = blocks[b];
        if (bl.struct_size != sizeof(sdpx_gpu_exact_block))
            return abi_invalid("block struct_size is not sizeof(sdpx_gpu_exact_block)");
        if (bl.rows && (!bl.y || !bl.d || (bl.ncols && !bl.columns))) return abi_invalid("null block array");
        for (uint64_t c = 0; c < bl.ncols; c++) {
            if (bl.columns[c] >= m) return abi_invalid("block column out of range");
            if (c && bl.columns[c] <= bl.columns[c - 1]) return abi_invalid("block columns must ascend");
        }
        K += bl.rows;
    }
    return 0;
}

// out (host records when !dev, else device; ldc = m): entries i <= j
int32_t blocks_upper_device(int limbs, uint64_t m, uint64_t nblocks, const sdpx_gpu_exact_block *blocks, void *out,
                            bool dev) {
    Device &dv = device();
    std::lock_guard<std::recursive_mutex> g(dv.mu);
    hipStream_t s = dv.stream;
    const size_t vb = value_bytes(limbs);
    uint64_t K = 0, vals = 0, ncol = 0, max_elems = 0;
    for (uint64_t b = 0; b < nblocks; b++) {
        K += blocks[b].rows;
        vals += blocks[b].rows * blocks[b].ncols + blocks[b].rows;
        ncol += blocks[b].ncols;
        max_elems = std::max<uint64_t>(max_elems, blocks[b].rows * blocks[b].ncols);
    }
    uint8_t *Z = (uint8_t *)dv.ws.get(GB_Z, std::max<uint64_t>(K * m, 1) * vb);
    uint8_t *Y = (uint8_t *)dv.ws.get(GB_Y, std::max<uint64_t>(K * m, 1) * vb);
    uint8_t *V = dev ? nullptr : (uint8_t *)dv.ws.get(GB_VALS, std::max<uint64_t>(vals, 1) * vb);
    int64_t *C = (int64_t *)dv.ws.get(GB_COLS, sizeof(int64_t) * std::max<uint64_t>(ncol, 1));
    BlockScatter *D = (BlockScatter *)dv.ws.get(GB_DESC, sizeof(BlockScatter) * std::max<uint64_t>(nblocks, 1));
    std::vector<int64_t> cols(ncol);
    std::vector<BlockScatter> desc(nblocks);
    uint64_t vo = 0, co = 0, r0 = 0;
    for (uint64_t b = 0; b < nblocks; b++) {
        const auto &bl = blocks[b];
        const void *y = bl.y, *d = bl.d;
        if (!dev) {  // stage the block's values
            if (bl.rows * bl.ncols) upload_async(V + vo * vb, bl.y, bl.rows * bl.ncols * vb);
            y = V + vo * vb;
            vo += bl.rows * bl.ncols;
            if (bl.rows) upload_async(V + vo * vb, bl.d, bl.rows * vb);
            d = V + vo * vb;
            vo += bl.rows;
        }
        for (uint64_t c = 0; c < bl.ncols; c++) cols[co + c] = (int64_t)bl.columns[c];
        desc[b] = BlockScatter{y, d, (int64_t)bl.rows, (int64_t)bl.ncols, (int64_t)r0, C + co};
        co += bl.ncols;
        r0 += bl.rows;
    }
    if (ncol) upload_async(C, cols.data(), sizeof(int64_t) * ncol);
    if (nblocks) upload_async(D, desc.data(), sizeof(BlockScatter) * nblocks);
    mp_zero_values(limbs, Z, (int64_t)(K * m), s);
    mp_zero_values(limbs, Y, (int64_t)(K * m), s);
    mp_block_scatter(limbs, D, (int)nblocks, (int64_t)max_elems, Z, Y, (int64_t)K, s);
    uint8_t *O = dev ? (uint8_t *)out : (uint8_t *)dv.ws.get(GB_OUT, m * m * vb);
    if (!dev) upload_async(O, out, m * m * vb);  // the strict lower triangle stays as given
    RnsGemmArgs a{limbs, 'T', 'N', (int64_t)m, (int64_t)m, (int64_t)K, Z, (int64_t)std::max<uint64_t>(K, 1), Y,
                  (int64_t)std::max<uint64_t>(K, 1), true, (uint64_t)(m * m), O};
    const int r = rns_gemm_device(a);
    if (r != SDPX_GPU_OK) return r;
    if (!dev) HIP_CHECK(hipMemcpyAsync(out, O, m * m * vb, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return SDPX_GPU_OK;
}

}  // namespace sdpx

extern "C" int32_t sdpx_gpu_rns_gemm_blocks_upper_dev(uint32_t limbs, uint64_t m, uint64_t nblocks,
                                                      const sdpx_gpu_exact_block *blocks, void *out) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        uint64_t K = 0;
        if (int32_t r = sdpx::check_exact_blocks(m, nblocks, blocks, K)) return r;
        if (m == 0) return SDPX_GPU_OK;
        if (!out) return sdpx::abi_invalid("null output");
        return sdpx::blocks_upper_device((int)limbs, m, nblocks, blocks, out, true);
    });
}
