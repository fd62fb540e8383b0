// Bimarginal counts as an int8 one-hot Gram product on the matrix cores.
//
// With X[n][r] = [s_{n, r/q} == r % q] (N x L*q, one-hot, never stored),
// count_{ij}(a, b) = (X^T X)[i*q + a][j*q + b] for i < j.  Each wave owns a
// 128 x 128 block of the flattened L*q x L*q Gram matrix (4 x 4 tiles of
// v_mfma_i32_32x32x32_i8), builds its int8 operands on the fly from the
// site-major residue bytes (16 sequences per lane per k-step, read from an
// LDS copy of the block's sites that LDS DMA refills 4 to 16 k-steps at a
// time, one load per site and 16 bytes instead of one per row), contracts
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

// Staging of residue bytes: STAGE k-steps (32 sequences each, so
// SB = 32 * STAGE bytes per site) of the sites a block's 128 rows and 128
// columns touch (at most 128/q + 2 each), double-buffered.  One LDS DMA
// instruction moves 16 bytes per lane, i.e. SPI = 1024 / SB sites.  Longer
// stages wait less often on their copies (16 k-steps: 10.9 ms, 8: 11.1,
// 4: 11.5, 2: 12.4 at 2^20 sequences, L=232 q=21); the host takes the
// longest whose double buffer fits 40 KiB, so that four waves share a CU.
template <uint32_t STAGE>
struct BcStage {
    static constexpr uint32_t SB = 32 * STAGE;
    static constexpr uint32_t LPS = SB / 16;   // lanes per site
    static constexpr uint32_t SPI = 64 / LPS;  // sites per instruction
    static_assert(SPI >= 1 && SPI * LPS == 64, "stage size");
};
inline uint32_t bc_max_sites(uint32_t q) { return 128 / q + 2; }
inline size_t bc_lds_bytes(uint32_t q, uint32_t stage) {
    const uint32_t spi = 32 / stage;
    return 2ull * ((2 * bc_max_sites(q) + spi - 1) / spi) * 1024;
}
inline uint32_t bc_stage_for(uint32_t q) {
    for (uint32_t st : {16u, 8u})
        if (bc_lds_bytes(q, st) <= 40960) return st;
    return 4;
}
typedef __attribute__((address_space(3))) void bc_lds_void;

// s_waitcnt vmcnt(0): every vector-memory operation of this wave, LDS DMA
// included, has completed.
__device__ __forceinline__ void wait_vmem_bc() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <uint32_t STAGE>
__global__ __launch_bounds__(64) void k_bicounts_mfma128(const uint8_t* __restrict__ seq,
                                                         uint32_t ldseq, uint32_t nseq,
                                                         uint32_t L, uint32_t q, uint32_t nblk,
                                                         uint32_t ksplit,
                                                         uint32_t* __restrict__ counts) {
    extern __shared__ __attribute__((aligned(16))) uint8_t bc_lds[];
    constexpr uint32_t BC_STAGE = STAGE, BC_SB = BcStage<STAGE>::SB,
                       BC_LPS = BcStage<STAGE>::LPS, BC_SPI = BcStage<STAGE>::SPI;
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

    // the sites of this block's rows (A) and columns (B), staged as
    // [A sites..., B sites...] x BC_SB bytes per buffer
    const uint32_t rA0 = rb * 128, rA1 = (rA0 + 127 < Lq) ? rA0 + 127 : Lq - 1;
    const uint32_t cB0 = cb * 128, cB1 = (cB0 + 127 < Lq) ? cB0 + 127 : Lq - 1;
    const uint32_t siA0 = rA0 / q, nsA = rA1 / q - siA0 + 1;
    const uint32_t siB0 = cB0 / q, nsB = cB1 / q - siB0 + 1;
    const uint32_t ninst = (nsA + nsB + BC_SPI - 1) / BC_SPI;
    const uint32_t bufbytes = ninst * 1024;

    uint32_t ldsA[4], a4A[4], naA[4], siteA[4], ldsB[4], a4B[4], naB[4], siteB[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const uint32_t r = rA0 + x * 32 + lr;
        const uint32_t si = r < Lq ? r / q : siA0;
        siteA[x] = si;
        a4A[x] = r < Lq ? (r - si * q) * 0x01010101u : 0x7f7f7f7fu;
        naA[x] = ~a4A[x];
        ldsA[x] = (si - siA0) * BC_SB + 16 * g;
        const uint32_t c = cB0 + x * 32 + lr;
        const uint32_t sj = c < Lq ? c / q : siB0;
        siteB[x] = sj;
        a4B[x] = c < Lq ? (c - sj * q) * 0x01010101u : 0x7f7f7f7fu;
        naB[x] = ~a4B[x];
        ldsB[x] = (nsA + sj - siB0) * BC_SB + 16 * g;
    }
    v16i acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[x][y][e] = 0;

    // LDS DMA of one stage: instruction k moves sites BC_SPI*k onward (lane
    // l: site BC_SPI*k + l/BC_LPS, 16-byte piece l%BC_LPS of the stage's
    // 32*BC_STAGE sequences);
    // slots past the block's sites load a valid site that nothing reads.
    // The last stage of a range can reach past the last step (and, for the
    // last site, past the buffer): such pieces are never read, and their
    // source offsets are clamped into the buffer.
    const uint32_t last16 = ldseq * L - 16;
    auto stage = [&](uint32_t buf, uint32_t step) {
        const uint32_t n0 = step * 32;
        for (uint32_t k = 0; k < ninst; k++) {
            const uint32_t ts = BC_SPI * k + l / BC_LPS;
            const uint32_t site = ts < nsA ? siA0 + ts : (ts < nsA + nsB ? siB0 + (ts - nsA) : siA0);
            const uint32_t off = site * ldseq + n0 + 16 * (l % BC_LPS);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rs, (bc_lds_void*)(bc_lds + buf * bufbytes + k * 1024), 16,
                off < last16 ? off : last16, 0, 0, 0);
        }
    };
    auto operands = [&](const uint8_t* sbase, uint32_t ks, v4i (&A)[4], v4i (&B)[4]) {
        v4i ra[4], rbv[4];
#pragma unroll
        for (int x = 0; x < 4; x++) {
            ra[x] = *reinterpret_cast<const v4i*>(sbase + ldsA[x] + 32 * ks);
            rbv[x] = *reinterpret_cast<const v4i*>(sbase + ldsB[x] + 32 * ks);
        }
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
    // full steps (all 32 sequences < nseq), staged BC_STAGE steps at a time
    // one stage ahead; the ragged last step, if any, goes byte by byte
    const uint32_t full_end = nseq / 32;
    const uint32_t e1 = s1 < full_end ? s1 : full_end;
    if (s0 < e1) {
        stage(0, s0);
        uint32_t buf = 0;
        for (uint32_t st0 = s0; st0 < e1; st0 += BC_STAGE) {
            wait_vmem_bc();
            __syncthreads();   // this stage's bytes are in LDS
            if (st0 + BC_STAGE < e1) stage(buf ^ 1, st0 + BC_STAGE);
            const uint8_t* sbase = bc_lds + buf * bufbytes;
            // one step at a time (a guard per step also keeps the compiler
            // from hoisting all the stage's LDS reads, which would spill)
#pragma unroll
            for (uint32_t ks = 0; ks < BC_STAGE; ks++) {
                if (st0 + ks < e1) {
                    v4i A[4], B[4];
                    operands(sbase, ks, A, B);
                    products(A, B);
                }
            }
            __syncthreads();   // reads of this buffer done before it is refilled
            buf ^= 1;
        }
    }
    if (s1 > full_end && full_end >= s0) {
        // ragged tail: the last, partial step of this part (bytes past nseq
        // count for nothing)
        const uint32_t n = full_end * 32 + 16 * g;
        v4i A[4], B[4];
#pragma unroll
        for (int x = 0; x < 4; x++) {
            v4i ra = onehot16(seq, ldseq, siteA[x], a4A[x], a4A[x] != 0x7f7f7f7fu, n, nseq);
            v4i rbv = onehot16(seq, ldseq, siteB[x], a4B[x], a4B[x] != 0x7f7f7f7fu, n, nseq);
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
