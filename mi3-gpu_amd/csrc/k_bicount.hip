// Bimarginal counts as an int8 one-hot Gram product on the matrix cores.
//
// With X[n][r] = [s_{n, r/q} == r % q] (N x L*q, one-hot, never stored),
// count_{ij}(a, b) = (X^T X)[i*q + a][j*q + b] for i < j.  Each wave owns a
// 64 x 64 block of the flattened L*q x L*q Gram matrix (2 x 2 tiles of
// v_mfma_i32_32x32x32_i8), builds its int8 operands on the fly from the
// site-major residue bytes (16 sequences per lane per k-step), contracts
// over all sequences with int32 accumulation, and scatters the i < j entries
// into the packed (L(L-1)/2, q*q) uint32 table.  Integer accumulation makes
// the result exact whatever the summation order.
// Replaces countBivariate (mcmc.cl:464-516); matches np.bincount per pair
// (getMarginals.py:80-86).
#include "common.h"

namespace mi3 {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// bytes of `v` equal to `a` -> 0x01, others -> 0x00
__device__ __forceinline__ uint32_t eq_bytes(uint32_t v, uint32_t a4) {
    uint32_t x = v ^ a4;
    uint32_t t = ((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x;
    return (~t & 0x80808080u) >> 7;
}

// One-hot operand of row r (site r/q, residue r%q) for 16 sequences
// starting at n: 16 int8 values in 4 dwords.
__device__ __forceinline__ v4i onehot16(const uint8_t* __restrict__ seq, uint32_t ldseq,
                                        uint32_t site, uint32_t a4, bool valid, uint32_t n,
                                        uint32_t nseq) {
    v4i r = {0, 0, 0, 0};
    if (valid) {
        const uint8_t* p = seq + (size_t)site * ldseq + n;
        uint4 v;
        if (n + 16 <= nseq) {
            v = *reinterpret_cast<const uint4*>(p);
            r[0] = (int)eq_bytes(v.x, a4);
            r[1] = (int)eq_bytes(v.y, a4);
            r[2] = (int)eq_bytes(v.z, a4);
            r[3] = (int)eq_bytes(v.w, a4);
        } else {
            // ragged tail: bytes past nseq contribute nothing
            uint32_t w[4] = {0, 0, 0, 0};
            for (uint32_t e = 0; e < 16; e++) {
                if (n + e < nseq && p[e] == (uint8_t)(a4 & 0xff)) w[e >> 2] |= 1u << (8 * (e & 3));
            }
            r[0] = (int)w[0]; r[1] = (int)w[1]; r[2] = (int)w[2]; r[3] = (int)w[3];
        }
    }
    return r;
}

// grid: one wave per (row block rb, col block cb) with rb <= cb, blocks of 64.
__global__ __launch_bounds__(64) void k_bicounts_mfma(const uint8_t* __restrict__ seq,
                                                      uint32_t ldseq, uint32_t nseq, uint32_t L,
                                                      uint32_t q, uint32_t nblk,
                                                      uint32_t* __restrict__ counts) {
    // decode the triangular block index
    uint32_t t = blockIdx.x;
    uint32_t rb = 0;
    while (t >= nblk - rb) {
        t -= nblk - rb;
        rb++;
    }
    const uint32_t cb = rb + t;
    const uint32_t l = threadIdx.x;
    const uint32_t lr = l & 31, g = l >> 5;
    const uint32_t Lq = L * q;

    // operand rows of this lane: two row tiles (A) and two column tiles (B)
    uint32_t siteA[2], a4A[2], siteB[2], a4B[2];
    bool vA[2], vB[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
        uint32_t r = rb * 64 + x * 32 + lr;
        vA[x] = r < Lq;
        siteA[x] = vA[x] ? r / q : 0;
        a4A[x] = (r - siteA[x] * q) * 0x01010101u;
        uint32_t c = cb * 64 + x * 32 + lr;
        vB[x] = c < Lq;
        siteB[x] = vB[x] ? c / q : 0;
        a4B[x] = (c - siteB[x] * q) * 0x01010101u;
    }
    v16i acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[x][y][e] = 0;

    for (uint32_t n0 = 0; n0 < nseq; n0 += 32) {
        const uint32_t n = n0 + 16 * g;
        v4i A0 = onehot16(seq, ldseq, siteA[0], a4A[0], vA[0], n, nseq);
        v4i A1 = onehot16(seq, ldseq, siteA[1], a4A[1], vA[1], n, nseq);
        v4i B0 = onehot16(seq, ldseq, siteB[0], a4B[0], vB[0], n, nseq);
        v4i B1 = onehot16(seq, ldseq, siteB[1], a4B[1], vB[1], n, nseq);
        acc[0][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B1, acc[1][1], 0, 0, 0);
    }

    // D layout of a 32x32 MFMA tile: value e of lane l is row
    // 8*(e/4) + 4*(l/32) + e%4, column l%32.
    const uint32_t qq = q * q;
#pragma unroll
    for (int x = 0; x < 2; x++) {
#pragma unroll
        for (int y = 0; y < 2; y++) {
            uint32_t c = cb * 64 + y * 32 + lr;
            if (c >= Lq) continue;
            uint32_t jsite = c / q, b = c - jsite * q;
#pragma unroll
            for (int e = 0; e < 16; e++) {
                uint32_t r = rb * 64 + x * 32 + 8 * (e / 4) + 4 * g + (e % 4);
                if (r >= Lq) continue;
                uint32_t isite = r / q, a = r - isite * q;
                if (isite < jsite)
                    counts[(size_t)pair_index(isite, jsite, L) * qq + a * q + b] = (uint32_t)acc[x][y][e];
            }
        }
    }
}

}  // namespace mi3
