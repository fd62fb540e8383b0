// Bimarginal counts as an int8 one-hot Gram product on the matrix cores.
//
// With X[n][r] = [s_{n, r/q} == r % q] (N x L*q, one-hot, never stored),
// count_{ij}(a, b) = (X^T X)[i*q + a][j*q + b] for i < j.  Each wave owns a
// 128 x 128 block of the flattened L*q x L*q Gram matrix (4 x 4 tiles of
// v_mfma_i32_32x32x32_i8), builds its int8 operands on the fly from the
// site-major residue bytes (16 sequences per lane per k-step), contracts
// over its range of sequences with int32 accumulation, and scatters the
// i < j entries into the packed (L(L-1)/2, q*q) uint32 table.  Integer accumulation makes
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

// grid: ksplit waves per (row block rb, col block cb), rb <= cb, blocks of
// 128 (256 int32 accumulators per lane, held in AGPRs), so every one-hot
// operand built from a 16-byte load feeds 4 matrix instructions.  The
// sequence range is split over the ksplit waves of a block, whose partial
// counts are added with integer atomics (exact in any order; the table must
// be zeroed first).  Residue bytes are loaded two k-steps ahead of use.
//
// Operands: row operands are -128 (0x80) where the residue matches, column
// operands +1, so each co-occurrence adds -128 and count = -acc / 128
// (|acc| <= 128 * nseq < 2^31 for nseq < 2^24).  With residues < 32 every
// byte of x = v ^ a is below 0x80, so 0x80808080 - x borrows nothing across
// bytes and its top bit per byte is set exactly where v == a; it is formed
// in one v_xad_u32 as (v ^ ~a) + 0x80808081.  Rows past L*q compare against
// 0x7f, which no residue equals.
__device__ __forceinline__ uint32_t match_hi(uint32_t v, uint32_t na4) {
    return (v ^ na4) + 0x80808081u;
}

__global__ __launch_bounds__(64) void k_bicounts_mfma128(const uint8_t* __restrict__ seq,
                                                         uint32_t ldseq, uint32_t nseq,
                                                         uint32_t L, uint32_t q, uint32_t nblk,
                                                         uint32_t ksplit,
                                                         uint32_t* __restrict__ counts) {
    uint32_t t = blockIdx.x / ksplit;
    const uint32_t part = blockIdx.x - t * ksplit;
    uint32_t rb = 0;
    while (t >= nblk - rb) {
        t -= nblk - rb;
        rb++;
    }
    const uint32_t cb = rb + t;
    const uint32_t l = threadIdx.x;
    const uint32_t lr = l & 31, g = l >> 5;
    const uint32_t Lq = L * q;
    // sequence range of this part, in steps of 32
    const uint32_t nsteps = (nseq + 31) / 32;
    const uint32_t s0 = (uint32_t)(((uint64_t)nsteps * part) / ksplit);
    const uint32_t s1 = (uint32_t)(((uint64_t)nsteps * (part + 1)) / ksplit);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(seq), (short)0,
                                          (int)(ldseq * L), 0x00020000);

    uint32_t offA[4], a4A[4], naA[4], offB[4], a4B[4], naB[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const uint32_t r = rb * 128 + x * 32 + lr;
        const uint32_t si = r < Lq ? r / q : 0;
        a4A[x] = r < Lq ? (r - si * q) * 0x01010101u : 0x7f7f7f7fu;
        naA[x] = ~a4A[x];
        offA[x] = si * ldseq + 16 * g;
        const uint32_t c = cb * 128 + x * 32 + lr;
        const uint32_t sj = c < Lq ? c / q : 0;
        a4B[x] = c < Lq ? (c - sj * q) * 0x01010101u : 0x7f7f7f7fu;
        naB[x] = ~a4B[x];
        offB[x] = sj * ldseq + 16 * g;
    }
    v16i acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[x][y][e] = 0;

    auto load = [&](uint32_t step, v4i (&ra)[4], v4i (&rbv)[4]) {
        const uint32_t n = step * 32;   // wave-uniform: the scalar offset
#pragma unroll
        for (int x = 0; x < 4; x++) {
            ra[x] = __builtin_amdgcn_raw_buffer_load_b128(rs, offA[x], n, 0);
            rbv[x] = __builtin_amdgcn_raw_buffer_load_b128(rs, offB[x], n, 0);
        }
    };
    auto operands = [&](const v4i (&ra)[4], const v4i (&rbv)[4], v4i (&A)[4], v4i (&B)[4]) {
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                A[x][e] = (int)(match_hi((uint32_t)ra[x][e], naA[x]) & 0x80808080u);
                B[x][e] = (int)((match_hi((uint32_t)rbv[x][e], naB[x]) >> 7) & 0x01010101u);
            }
    };
    auto products = [&](const v4i (&A)[4], const v4i (&B)[4]) {
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++)
                acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[x], B[y], acc[x][y], 0, 0, 0);
    };
    // full steps (all 32 sequences < nseq), two buffers of raw bytes loaded
    // two steps ahead; the ragged last step, if any, goes byte by byte
    const uint32_t full_end = nseq / 32;
    const uint32_t e1 = s1 < full_end ? s1 : full_end;
    v4i r0a[4], r0b[4], r1a[4], r1b[4], A[4], B[4];
    if (s0 < e1) load(s0, r0a, r0b);
    if (s0 + 1 < e1) load(s0 + 1, r1a, r1b);
    for (uint32_t st = s0; st < e1; st += 2) {
        operands(r0a, r0b, A, B);
        if (st + 2 < e1) load(st + 2, r0a, r0b);
        products(A, B);
        if (st + 1 < e1) {
            operands(r1a, r1b, A, B);
            if (st + 3 < e1) load(st + 3, r1a, r1b);
            products(A, B);
        }
    }
    if (s1 > full_end && full_end >= s0) {
        // ragged tail: the last, partial step of this part (bytes past nseq
        // count for nothing)
        const uint32_t n = full_end * 32 + 16 * g;
#pragma unroll
        for (int x = 0; x < 4; x++) {
            v4i ra = onehot16(seq, ldseq, offA[x] / ldseq, a4A[x], a4A[x] != 0x7f7f7f7fu, n, nseq);
            v4i rbv = onehot16(seq, ldseq, offB[x] / ldseq, a4B[x], a4B[x] != 0x7f7f7f7fu, n, nseq);
            // onehot16 gives 0/1 bytes; rows take -128
#pragma unroll
            for (int e = 0; e < 4; e++) ra[e] = (int)((uint32_t)ra[e] << 7);
            A[x] = ra;
            B[x] = rbv;
        }
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++)
                acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[x], B[y], acc[x][y], 0, 0, 0);
    }

    const uint32_t qq = q * q;
#pragma unroll
    for (int x = 0; x < 4; x++) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const uint32_t c = cb * 128 + y * 32 + lr;
            if (c >= Lq) continue;
            const uint32_t jsite = c / q, b = c - jsite * q;
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const uint32_t r = rb * 128 + x * 32 + 8 * (e / 4) + 4 * g + (e % 4);
                if (r >= Lq) continue;
                const uint32_t isite = r / q, a = r - isite * q;
                if (isite < jsite) {
                    const uint32_t cnt = (0u - (uint32_t)acc[x][y][e]) >> 7;
                    uint32_t* dst = counts + (size_t)pair_index(isite, jsite, L) * qq + a * q + b;
                    if (ksplit == 1)
                        *dst = cnt;
                    else
                        atomicAdd(dst, cnt);
                }
            }
        }
    }
}

}  // namespace mi3
