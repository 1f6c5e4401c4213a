// Modulus plans of the exact residue (RNS) product on INT8 matrix cores.
//
// A plan is a prefix of one fixed, ordered modulus list:
//   * single-digit moduli: the largest prime power <= 256 of every prime
//     (256, 251, 243, 241, ..., 17), descending; their symmetric residues fit
//     one signed byte, so a product costs one INT8 GEMM per modulus;
//   * two-digit moduli: primes in (256, 65279], descending; a symmetric
//     residue r (|r| <= 32639) is h*256 + l with h, l signed bytes, so a
//     product costs four INT8 GEMMs (hh, hl + lh, ll) per modulus.
// Single-digit moduli carry >= 4.09 bits per GEMM, two-digit ones ~4, so the
// list is ordered by bits per GEMM and every plan takes its cheapest prefix.
// Every INT8 product term is at most 2^14 in magnitude, so INT32
// accumulation is exact for inner dimensions below 2^17.
//
// The role of the plan matches the reference's `Plan`
// (crates/solver/src/algebra/dense/blas/rns_blas.rs:705-756): enough
// moduli that M = prod(q) exceeds 4|S| for every exact dot S, so the signed
// CRT lift recovers S exactly (rns_blas.rs:2202-2205).  The moduli, digit
// split and CRT form are this library's own.
#pragma once
#include <cstdint>
#include <vector>

namespace sdpx {

struct RnsPlan {
    int count = 0;         // moduli L
    int single = 0;        // of which the first `single` are one-digit
    int planes = 0;        // INT8 planes per operand: single + 2*(count-single)
    double bits = 0;       // log2(M)
    std::vector<uint32_t> q;       // moduli
    std::vector<uint32_t> inv;     // (M/q_l)^-1 mod q_l
    std::vector<double> qinv;      // 1/q_l
    // CRT constants in 24-bit limbs, least significant first, `limbs` each.
    int limbs = 0;
    std::vector<double> c;         // count x limbs: M/q_l
    std::vector<double> m;         // limbs: M
    // Matrix-core CRT constants.  V = sum_l x'_l (M/q_l) for balanced-digit
    // x'_l = 256 hi + lo is the INT8 product of the digit pairs with table
    // [t][2l] = c_{l,t}, [t][2l+1] = c_{l,t-1}, c_{l,t} the balanced base-256
    // digits of M/q_l (t <= crt_jb); rows crt_nt = crt_jb + 2, padded to
    // crt_npad rows of 2 crt_lpad bytes.  M in crt_nu 32-bit words (room for
    // |V| <= (count/2) M and a sign), m_top = M / 2^(32 max(crt_nu - 3, 0))
    // (M's top min(crt_nu, 3) words).
    int crt_jb = 0, crt_nt = 0, crt_npad = 0, crt_lpad = 0, crt_nu = 0;
    double m_top = 0;
    std::vector<int8_t> crt_table;
    std::vector<uint32_t> m32;
};

// All moduli in plan order.
const std::vector<uint32_t> &rns_moduli();
// Number of single-digit moduli at the head of rns_moduli().
int rns_single_count();
// The cheapest prefix plan with log2(M) >= needed_bits, or nullptr when the
// modulus list cannot reach it.  Plans are cached and never freed.
const RnsPlan *rns_plan_for_bits(double needed_bits);
// Bits the product of op(A) and op(B) needs: operand integer widths wa, wb
// (|a'| < 2^wa), inner dimension k; M > 4 k 2^(wa+wb) with one spare bit.
double rns_needed_bits(int64_t wa, int64_t wb, int64_t k);

}  // namespace sdpx
