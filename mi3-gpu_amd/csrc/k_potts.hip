// Metropolis sweep and batch Potts energies for gfx950.
//
// Both kernels use one work distribution:
//   * a walker (or sequence) is owned by a half-wave: lane j (0..31) holds the
//     residues of sites m = j + 32k, k = 0..4*KW-1, packed 4 per u32 register,
//     for S walkers ("slots"); a wave therefore owns 2*S walkers and a
//     work-group of NWAVES waves owns W = 2*S*NWAVES walkers for the whole
//     launch (the walker state never leaves registers between site updates);
//   * couplings for one site i are streamed through LDS in chunks of 32
//     neighbour sites, laid out [a*q + c][32] so that the 32 lanes of a
//     half-wave (32 distinct neighbour sites) gather from 32 distinct banks
//     whatever residues their walkers hold;
//   * chunks are double-buffered in LDS: the next chunk is copied from global
//     memory straight into LDS (buffer_load ... lds) while the current one is
//     gathered, one barrier per chunk.
//
// Chunked coupling layout (built by k_unpack_chunked in k_misc.hip):
//   Ju[i][c][a*q + b][jl] = J_{i,m}(a, b),  m = 32c + jl,
// symmetric (both i<m and i>m filled), zero for m == i and m >= L.
#include "common.h"
#include "launch.h"

namespace mi3 {

// Native 4-float vector (HIP's float4 is a union-based struct, which keeps
// arrays of it out of registers).
typedef float f4v __attribute__((ext_vector_type(4)));

// w[idx] for a wave-uniform idx, as a select chain over registers.  The
// empty asm keeps the compiler from turning the chain back into a dynamic
// index (which would move the state array to scratch memory).
template <int KW>
__device__ __forceinline__ uint32_t select_word(const uint32_t (&w)[KW], uint32_t idx) {
    uint32_t r = w[0];
#pragma unroll
    for (int k = 1; k < KW; k++) {
        uint32_t t = w[k];
        asm volatile("" : "+v"(t));
        r = (idx == (uint32_t)k) ? t : r;
    }
    return r;
}

// Word k of a walker's packed state.  Static indices address registers
// directly; for KW >= 3 the (unrolled) index goes through the opaque select
// chain so the state array stays in registers.
template <int KW>
__device__ __forceinline__ uint32_t state_word(const uint32_t (&w)[KW], uint32_t k) {
    if constexpr (KW <= 2)
        return (KW == 1 || k == 0) ? w[0] : w[KW - 1];
    else
        return select_word<KW>(w, k);
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Sweep gather schedule: slots per group and groups of reads in flight
// (8 and 1, 8 and 2, 4 and 2, 4 and 3, 16 and 1 time alike at L=232 q=21).
constexpr int SWEEP_G = 8;
constexpr int SWEEP_P = 1;

// Wave-wide lane masks: bit `lane` set when (lane & m) != 0.
__device__ __forceinline__ constexpr uint64_t lane_bit_mask(int m) {
    return m == 16 ? 0xFFFF0000FFFF0000ull
         : m == 8  ? 0xFF00FF00FF00FF00ull
         : m == 4  ? 0xF0F0F0F0F0F0F0F0ull
         : m == 2  ? 0xCCCCCCCCCCCCCCCCull
                   : 0xAAAAAAAAAAAAAAAAull;
}

// mask[lane] ? a : b as one v_cndmask.  Written as asm so that the
// optimizer cannot turn a select between two elements of one register array
// back into a lane-varying array index (which it lowers to a long
// compare/select chain).
__device__ __forceinline__ float lane_select(uint64_t mask, float a, float b) {
    float r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(mask));
    return r;
}

// Value of v in the partner lane of a reduction level below 16, within a
// row of 16 lanes, by DPP: m=8 row_ror:8 (lane ^ 8), m=4 row_half_mirror
// (lane ^ 7: the pair differs in bit 2, and in bits 0-1, which the later
// levels sum over anyway), m=2 and m=1 quad permutations (lane ^ 2, ^ 1).
__device__ __forceinline__ float dpp_partner(float v, int m) {
    const int x = __float_as_int(v);
    switch (m) {
        case 8: return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false));
        case 4: return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));
        case 2: return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
        default: return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    }
}

// Reduce S per-lane partial sums over the 32 lanes of each half-wave at once
// (butterfly with operand exchange): afterwards lane j holds the full sum of
// slot j >> (5 - log2 S).  The first level pairs lanes j and j + 16: one
// v_permlane16_swap exchanges the odd 16-lane rows of v[s] with the even
// rows of v[s + h], so one add leaves rows 0 and 2 with slot s and rows 1
// and 3 with slot s + h, with no selects.  Each later level keeps one of two
// slots by the lane's bit m (v_cndmask with constant lane masks) and adds
// the other slot's value from the partner lane through DPP.  No LDS.
template <int S>
__device__ __forceinline__ float transposed_sum(float (&v)[S]) {
    int n = S;
    if (n > 1) {
        const int h = n / 2;
#pragma unroll
        for (int s = 0; s < S / 2; s++) {
            if (s < h) {
                const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[s]),
                                                                __float_as_uint(v[s + h]), false,
                                                                false);
                v[s] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
            }
        }
        n = h;
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[0]),
                                                        __float_as_uint(v[0]), false, false);
        v[0] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int l = 1; l < 5; l++) {
        const int m = 16 >> l;
        if (n > 1) {
            const int h = n / 2;
            const uint64_t mk = lane_bit_mask(m);
#pragma unroll
            for (int s = 0; s < S / 2; s++) {
                if (s < h) {
                    float send = lane_select(mk, v[s], v[s + h]);
                    float keep = lane_select(mk, v[s + h], v[s]);
                    v[s] = keep + dpp_partner(send, m);
                }
            }
            n = h;
        } else {
            v[0] += dpp_partner(v[0], m);
        }
    }
    return v[0];
}

// Buffer resource for raw buffer loads (gfx9 dword3 = 0x00020000): 32-bit
// offsets, so staging addresses need no 64-bit arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}

// Copy one padded chunk (PF * NT float4) from global memory straight into
// LDS with buffer_load ... lds: lane l of wave w moves float4 number
// k*NT + 64*w + l; the hardware writes the 64 float4 of a wave contiguously
// at the wave-uniform LDS address.  No registers hold the data in flight;
// completion is awaited with s_waitcnt vmcnt(0) before the next barrier.
typedef __attribute__((address_space(3))) void lds_void;

template <int PF, int NT>
__device__ __forceinline__ void chunk_dma(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                          char* lds_buf, int wave) {
#pragma unroll
    for (int k = 0; k < PF; k++)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            r, (lds_void*)(lds_buf + (k * NT + wave * 64) * 16), 16, voff, soff + k * NT * 16, 0,
            0);
}

// s_waitcnt vmcnt(0) (expcnt, lgkmcnt left at their maxima): every
// vector-memory operation of this wave, including LDS DMA, has completed.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// s_waitcnt vmcnt(N): at most N vector-memory operations of this wave are
// still in flight (gfx9 encoding: vmcnt bits 3:0 and 15:14).  Loads complete
// in issue order, so all but the youngest N have landed.
template <int N>
__device__ __forceinline__ void wait_vmem_upto() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 15) | ((N >> 4) << 14));
}

// Work-group barrier that leaves LDS DMA in flight.  __syncthreads() waits
// for vmcnt(0) (its workgroup release fence covers LDS DMA writes), which
// drains every copy issued ahead; here only this wave's LDS and scalar loads
// are awaited (lgkmcnt(0) as a builtin, so that the compiler's own wait
// tracking sees them complete), then s_barrier, whose memory clobber keeps
// memory accesses on their side.
__device__ __forceinline__ void barrier_copies_in_flight() {
    __builtin_amdgcn_s_waitcnt(0xC07F);
    asm volatile("s_barrier" ::: "memory");
}

// Value of `v` in lane (src + 32*half) of each half-wave, src a compile-time
// lane number < 32 (ds_swizzle bitmask mode: and 0, or src).
#define MI3_SWZ_CASE(n) \
    case n:             \
        return __builtin_amdgcn_ds_swizzle(v, (n) << 5);
__device__ __forceinline__ int half_bcast(int v, int src) {
    switch (src) {
        MI3_SWZ_CASE(0) MI3_SWZ_CASE(1) MI3_SWZ_CASE(2) MI3_SWZ_CASE(3) MI3_SWZ_CASE(4)
        MI3_SWZ_CASE(5) MI3_SWZ_CASE(6) MI3_SWZ_CASE(7) MI3_SWZ_CASE(8) MI3_SWZ_CASE(9)
        MI3_SWZ_CASE(10) MI3_SWZ_CASE(11) MI3_SWZ_CASE(12) MI3_SWZ_CASE(13) MI3_SWZ_CASE(14)
        MI3_SWZ_CASE(15) MI3_SWZ_CASE(16) MI3_SWZ_CASE(17) MI3_SWZ_CASE(18) MI3_SWZ_CASE(19)
        MI3_SWZ_CASE(20) MI3_SWZ_CASE(21) MI3_SWZ_CASE(22) MI3_SWZ_CASE(23) MI3_SWZ_CASE(24)
        MI3_SWZ_CASE(25) MI3_SWZ_CASE(26) MI3_SWZ_CASE(27) MI3_SWZ_CASE(28) MI3_SWZ_CASE(29)
        MI3_SWZ_CASE(30) default: return __builtin_amdgcn_ds_swizzle(v, 31 << 5);
    }
}
#undef MI3_SWZ_CASE

// ---------------------------------------------------------------------------
// Chunk ring with overlapping slots, for the 57,344-byte padded chunks of
// q = 17..21 (7 float4 per thread of 512 staging threads).  Two such chunks
// fill 112 KB of the 160 KB of LDS, and three do not fit, so a plain ring can
// only double-buffer: the copy of chunk n+1 starts when chunk n-1 has been
// read and must land within the gathers of chunk n.  Here four slots
// X0..X3 (chunk n in X[n mod 4]) overlap in pairs two apart, X0/X2 and
// X1/X3, by 32 KB:
//   X0 = [0, 56K)   X2 = [24K, 80K)   X1 = [80K, 136K)   X3 = [104K, 160K)
// (K = 1024 bytes); slots one apart are disjoint.  A chunk's copy is split
// into its part outside the partner slot (24 KB, "early") and its part
// inside it (32 KB, "late"):
//   * early(n+2) is issued at the start of chunk n: it overlaps none of the
//     slots of chunks n and n+1;
//   * late(n+1) is issued at the start of chunk n: it overlaps only the slot
//     of chunk n-1, which every wave finished reading before the barrier
//     that started chunk n.
// So a chunk's copy has two chunks' time for 24 KB and one for 32 KB
// (double buffering gives one for all 56 KB), and LDS reads never meet a
// copy in flight.  The energy kernel streams the q = 17..21 chunks of its
// four-state-word shape this way (for fewer state words, and in the sweep,
// which is bound by VALU issue and LDS traffic, the ring was slower).  Per chunk each of the 8 staging waves issues 4 late then
// 3 early one-KB copies; at the end of chunk n only early(n+2), the 3
// youngest, may still be in flight (vmcnt 3), and the barrier that follows
// must not wait for it (barrier_copies_in_flight).
struct OverlapRing {
    static constexpr uint32_t PIECE = 1024;   // one wave's copy: 64 x 16 bytes
    static constexpr uint32_t CHB = 57344;    // padded chunk: 56 pieces
    static constexpr int EARLY_OPS = 3, LATE_OPS = 4;  // per staging wave
    static constexpr uint32_t BYTES = 163840;
    // slot base (0, 80K, 24K, 104K for r = 0..3), first piece of the late
    // part (the upper 32 pieces of X0 and X1, the lower 32 of X2 and X3) and
    // of the early part, as arithmetic on r (scalar, no branches)
    __device__ static uint32_t base(uint32_t r) { return (r & 1u) * 81920u + (r >> 1) * 24576u; }
    __device__ static uint32_t late_first(uint32_t r) { return 24u - (r >> 1) * 24u; }
    __device__ static uint32_t early_first(uint32_t r) { return (r >> 1) * 32u; }

    // copy `nops` pieces starting at piece `first` of the chunk at byte
    // offset `off` of the chunked layout into ring slot r; wave w < 8 moves
    // pieces first + 8*o + w
    template <int NOPS>
    __device__ static void copy(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t off,
                                char* lds, uint32_t r, uint32_t first, int wave) {
        char* dst = lds + base(r) + (first + wave) * PIECE;
        const uint32_t src = off + first * PIECE;
#pragma unroll
        for (int o = 0; o < NOPS; o++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void*)(dst + o * 8 * PIECE), 16, voff,
                                                     src + o * 8 * PIECE, 0, 0);
    }
};
static_assert(OverlapRing::CHB == 7 * 512 * 16, "padded chunk of the PF = 7 class");
static_assert(OverlapRing::BYTES == 106496 + OverlapRing::CHB, "slots fill the LDS");

template <int KW>
__device__ __forceinline__ uint32_t select_word_opaque(const uint32_t (&w)[KW], uint32_t idx) {
    if constexpr (KW == 1) return w[0];
    if constexpr (KW == 2) return idx ? w[1] : w[0];
    uint32_t r = w[0];
#pragma unroll
    for (int k = 1; k < KW; k++) {
        uint32_t t = w[k];
        asm("" : "+v"(t));
        r = (idx == (uint32_t)k) ? t : r;
    }
    return r;
}

// 128 * byte e of w (e a compile-time constant after unrolling) in one
// VALU op: a shift with an SDWA byte select on its source.  Left to itself
// the compiler makes byte 0 a shift and a mask.
__device__ __forceinline__ uint32_t byte_x128(uint32_t w, int e) {
    uint32_t r;
    switch (e) {
        case 0:
            asm("v_lshlrev_b32_sdwa %0, 7, %1 dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:DWORD src1_sel:BYTE_0" : "=v"(r) : "v"(w));
            break;
        case 1:
            asm("v_lshlrev_b32_sdwa %0, 7, %1 dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:DWORD src1_sel:BYTE_1" : "=v"(r) : "v"(w));
            break;
        case 2:
            asm("v_lshlrev_b32_sdwa %0, 7, %1 dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:DWORD src1_sel:BYTE_2" : "=v"(r) : "v"(w));
            break;
        default:
            asm("v_lshlrev_b32_sdwa %0, 7, %1 dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:DWORD src1_sel:BYTE_3" : "=v"(r) : "v"(w));
            break;
    }
    return r;
}

// a * b + c for 24-bit a, b in one op (the compiler turns the add of
// disjoint bit ranges into an or, which then stays a separate instruction).
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ uint32_t lane_select_u(uint64_t mask, uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(mask));
    return r;
}

// ---------------------------------------------------------------------------
// Metropolis sweep: nsteps single-site updates per walker.
//
// Step t of the launch is global step g = step0 + t (step0 even).  All
// walkers update the same site pos_list[t].  Walker w proposes residue
// b = mulhi(x, q) and accepts iff exp(-beta*dE) > (y >> 8) * 2^-24, where
// (x, y) are words (0,1) or (2,3) of Philox(key, (g>>1, w, TAG_WALK)) for
// even or odd g.  dE = sum_{m != pos} J_{pos,m}(b, s_m) - J_{pos,m}(a, s_m).
// Restates metropolis/DeltaEnergy (mcmc.cl:360-458) with counter-based
// streams in place of MWC64X.
//
// Per step: lane j draws the random numbers of walker slot j >> SH; each
// lane then fetches, for every slot, the walker's current residue a (from
// the lane owning site pos, ds_bpermute) and proposal b (ds_swizzle
// broadcast), gathers J(a, s_m) and J(b, s_m) for its sites chunk by chunk,
// and one transposed butterfly leaves each lane with the dE of its own slot;
// it decides, and a ballot tells the owner lane of pos which walkers accepted.
//
// sweep_block runs the walkers of work-group tile `block` (W = 2*S*NWAVES
// walkers starting at walker block*W); k_sweep is one tile per work-group,
// k_sweep_batch a persistent loop over the tiles of many models.
template <int KW, int S, int NWAVES, int PF>
__device__ __forceinline__ void sweep_block(
    const float* __restrict__ Ju, uint8_t* __restrict__ seq, const uint32_t* __restrict__ pos_list,
    uint32_t ldseq, uint32_t L, uint32_t q, uint32_t nchunk, uint32_t nsteps, uint32_t k0,
    uint32_t k1, uint64_t step0, uint32_t walker_base, float nbeta, uint32_t block) {
    constexpr int NT = 64 * NWAVES;
    constexpr int K = 4 * KW;
    constexpr int SH = 5 - ilog2(S);  // lane j reduces and decides slot j >> SH
    static_assert((1 << ilog2(S)) == S && S <= 32, "S must be a power of two <= 32");
    extern __shared__ f4v lds4[];
    char* ldsb = reinterpret_cast<char*>(lds4);

    // the padded chunk is PF float4 per thread of the first DMA_T threads
    // (the chunk layout's unit); waves beyond them take no part in staging
    constexpr int DMA_T = NT < 512 ? NT : 512;
    constexpr uint32_t chf = PF * DMA_T * 4;  // padded chunk (floats)
    constexpr uint32_t chb = chf * 4;      // padded chunk (bytes)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t j = lane & 31;
    const bool upper = lane >= 32;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t wfirst = block * (2 * S * NWAVES) + (wave * 2 + (upper ? 1 : 0)) * S;
    const uint32_t myslot = j >> SH;
    const __amdgpu_buffer_rsrc_t rJ = make_rsrc(Ju, L * nchunk * chb);
    const uint32_t voff = tid * 16;
    // staging waves: each copies one float4 per lane per pass; a single-pass
    // chunk (PF = 1, q <= 8) needs only the first ceil(q*q/8) waves, whose
    // 1 KB pieces hold its q*q*32 floats (the rest is padding nobody reads:
    // 3-4% faster at q = 3..4 than copying all 8 KB)
    const bool stage = (PF == 1) ? (uint32_t)wave < (q * q + 7) / 8 : wave < DMA_T / 64;

    // walker state: residues of sites j + 32k
    uint32_t st[S][KW];
#pragma unroll
    for (int s = 0; s < S; s++) {
#pragma unroll
        for (int kw = 0; kw < KW; kw++) {
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                uint32_t m = j + CH * (4 * kw + e);
                uint32_t v = (m < L) ? seq[(size_t)m * ldseq + wfirst + s] : 0u;
                word |= v << (8 * e);
            }
            st[s][kw] = word;
        }
        asm volatile("" ::: "memory");
    }

    uint32_t baseA[S], baseB[S];  // LDS byte offsets of rows a and b (+ lane column)
    float acc[S];
    uint32_t Pe = 0, Po = 0;      // this lane's walker draws for steps 2h, 2h+1

    if (nsteps == 0) return;
    // site of the current step; the next one is read one step ahead (with
    // the prefetch of its first chunk), so no step waits on a scalar load
    // (readfirstlane: the site is wave-uniform whatever the compiler can prove
    // about the list's memory, so masks built from it stay in SGPRs)
    uint32_t pos = __builtin_amdgcn_readfirstlane(pos_list[0]);
    uint32_t pos_next = nsteps > 1 ? __builtin_amdgcn_readfirstlane(pos_list[1]) : 0u;
    if (stage) chunk_dma<PF, DMA_T>(rJ, voff, pos * nchunk * chb, ldsb, wave);
    wait_vmem();
    __syncthreads();
    uint32_t buf = 0;
    const uint32_t qch4 = q * CH * 4;
    const uint32_t j4 = j * 4;

    for (uint32_t t = 0; t < nsteps; t++) {
        const uint32_t pos_after =
            (t + 2 < nsteps) ? __builtin_amdgcn_readfirstlane(pos_list[t + 2]) : 0u;
        const uint32_t kp = pos >> 5, owner = pos & 31;
        const uint32_t kwp = kp >> 2, shp = (kp & 3) * 8;
        if ((t & 1) == 0) {
            const uint64_t h = (step0 + t) >> 1;
            uint4 r = philox4x32_10(make_uint4((uint32_t)h, (uint32_t)(h >> 32),
                                               walker_base + wfirst + myslot, TAG_WALK),
                                    k0, k1);
            Pe = ((r.y >> 8) << 8) | mulhi_range(r.x, q);
            Po = ((r.w >> 8) << 8) | mulhi_range(r.z, q);
        }
        const uint32_t Pcur = (t & 1) ? Po : Pe;
        const uint32_t bmine = Pcur & 0xffu;
        const int own_addr = (int)((owner + (lane & 32)) * 4);
        // the proposal as a row offset, scaled by its sender so that every
        // receiving lane adds only its column
        const int brow = (int)__umul24(bmine, qch4);
        // residues a (from the owner lane of pos) and proposals b of every
        // slot, fetched in groups ahead of their use
        constexpr int GB = (S < 8) ? S : 8;
#pragma unroll
        for (int s0 = 0; s0 < S; s0 += GB) {
            uint32_t a[GB], b[GB];
#pragma unroll
            for (int g = 0; g < GB; g++) {
                const int s = s0 + g;
                const uint32_t av = __builtin_amdgcn_ubfe(select_word_opaque<KW>(st[s], kwp), shp, 8);
                a[g] = (uint32_t)__builtin_amdgcn_ds_bpermute(own_addr, (int)av);
                b[g] = (uint32_t)half_bcast(brow, s << SH);
            }
#pragma unroll
            for (int g = 0; g < GB; g++) {
                baseA[s0 + g] = mad_u24(a[g], qch4, j4);
                baseB[s0 + g] = b[g] + j4;
            }
        }
        // stream the couplings of site pos, chunk by chunk
#pragma unroll
        for (int k = 0; k < K; k++) {
            // (chunk 0 always exists: nchunk >= 1)
            if (k == 0 || (uint32_t)k < nchunk) {
                const bool last_chunk = (uint32_t)k + 1 == nchunk;
                const bool has_next = !last_chunk || (t + 1 < nsteps);
                if (has_next) {
                    uint32_t npos = last_chunk ? pos_next : pos;
                    uint32_t nc = last_chunk ? 0u : (uint32_t)k + 1;
                    if (stage)
                        chunk_dma<PF, DMA_T>(rJ, voff, (npos * nchunk + nc) * chb,
                                             ldsb + (buf ^ 1) * chb, wave);
                }
                const char* cb = ldsb + buf * chb;
                // gathers in groups of G slots, P groups in flight: the reads
                // of group i are issued before the sums of group i - P + 1,
                // and scheduling barriers keep the compiler from interleaving
                // them slot by slot (which exposes the LDS latency per slot)
                constexpr int G = (S < SWEEP_G) ? S : SWEEP_G;
                constexpr int NG = S / G;
                constexpr int P = (NG < SWEEP_P) ? NG : SWEEP_P;
                float vA[S], vB[S];
#pragma unroll
                for (int i = 0; i < NG + P - 1; i++) {
                    if (i < NG) {
#pragma unroll
                        for (int g = 0; g < G; g++) {
                            const int s = i * G + g;
                            static_assert(CH * 4 == 128, "row stride of a chunk");
                            const uint32_t off =
                                byte_x128(state_word<KW>(st[s], (uint32_t)k >> 2), k & 3);
                            vA[s] = *reinterpret_cast<const float*>(cb + baseA[s] + off);
                            vB[s] = *reinterpret_cast<const float*>(cb + baseB[s] + off);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    if (i >= P - 1) {
#pragma unroll
                        for (int g = 0; g < G; g++) {
                            const int s = (i - P + 1) * G + g;
                            // the first chunk starts the sums
                            if (k == 0) acc[s] = vB[s] - vA[s];
                            else acc[s] += vB[s] - vA[s];
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                wait_vmem();
                __syncthreads();
                buf ^= 1;
            }
        }
        // accept / reject: this lane decides for walker slot `myslot`
        const float dE = transposed_sum<S>(acc);
        const float u = (float)(Pcur >> 8) * 0x1.0p-24f;
        const uint64_t amask = __ballot(__expf(nbeta * dE) > u);
        // the owner lanes of pos take the proposals their half-wave accepted:
        // all proposal broadcasts are issued together (one LDS wait), the
        // state word holding pos is chosen once per step, and each slot's
        // lane mask takes three scalar operations
        uint32_t bnew[S];
        const int bsh = (int)(bmine << shp);  // the proposal at its byte of the word
#pragma unroll
        for (int s = 0; s < S; s++) bnew[s] = (uint32_t)half_bcast(bsh, s << SH);
        const uint32_t keep = ~(0xffu << shp);
#pragma unroll
        for (int kw = 0; kw < KW; kw++) {
            if (kwp == (uint32_t)kw) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    // bits 0 and 32 of the shifted ballot are the two
                    // half-waves' decisions for slot s; moved to the owner lanes
                    const uint64_t m = ((amask >> (s << SH)) & 0x0000000100000001ull) << owner;
                    const uint32_t nw = (st[s][kw] & keep) | bnew[s];
                    st[s][kw] = lane_select_u(m, nw, st[s][kw]);
                }
            }
        }
        pos = pos_next;
        pos_next = pos_after;
    }

    // write back
#pragma unroll
    for (int s = 0; s < S; s++) {
#pragma unroll
        for (int kw = 0; kw < KW; kw++) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                uint32_t m = j + CH * (4 * kw + e);
                if (m < L) seq[(size_t)m * ldseq + wfirst + s] = (uint8_t)(st[s][kw] >> (8 * e));
            }
        }
    }
}


template <int KW, int S, int NWAVES, int PF>
__global__ __launch_bounds__(64 * NWAVES) __attribute__((amdgpu_waves_per_eu(NWAVES / 4, NWAVES / 4))) void k_sweep(
    const float* __restrict__ Ju, uint8_t* __restrict__ seq, const uint32_t* __restrict__ pos_list,
    uint32_t ldseq, uint32_t L, uint32_t q, uint32_t nchunk, uint32_t nsteps, uint32_t k0,
    uint32_t k1, uint64_t step0, uint32_t walker_base, float nbeta) {
    sweep_block<KW, S, NWAVES, PF>(Ju, seq, pos_list, ldseq, L, q, nchunk, nsteps, k0, k1, step0,
                                   walker_base, nbeta, blockIdx.x);
}

// Persistent sweep over the tiles of one (KW, PF) class of a batch of
// models: every work-group takes tiles from a shared counter until none is
// left (the counter is cleared before each launch).
template <int KW, int S, int NWAVES, int PF>
__global__ __launch_bounds__(64 * NWAVES) __attribute__((amdgpu_waves_per_eu(NWAVES / 4, NWAVES / 4))) void
k_sweep_batch(const BatchModel* __restrict__ models, const BatchTile* __restrict__ tiles,
              uint32_t ntiles, uint32_t* __restrict__ counter) {
    __shared__ uint32_t s_tile;
    for (;;) {
        __syncthreads();   // previous tile's LDS reads done; s_tile free
        if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_tile);  // uniform: scalar loads below
        if (t >= ntiles) break;
        const BatchTile tl = tiles[t];
        const BatchModel& m = models[tl.model];
        // descriptor fields are uniform; say so (the kernel's own stores to
        // walker memory keep the compiler from proving it)
        auto u32 = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane(v); };
        // (readfirstlane returns int: widen through uint32_t, or a low word
        // with bit 31 set would sign-extend over the high word)
        auto ptr = [](const void* p) {
            const uint64_t v = (uint64_t)p;
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
            return (uint64_t)lo | ((uint64_t)hi << 32);
        };
        const uint64_t step0 = ptr((const void*)*m.step);
        sweep_block<KW, S, NWAVES, PF>(
            (const float*)ptr(m.Ju), (uint8_t*)ptr(m.seq), (const uint32_t*)ptr(m.pos), u32(m.ld),
            u32(m.L), u32(m.q), u32(m.nchunk), u32(m.nsteps), u32(m.k0), u32(m.k1), step0,
            u32(m.walker_base), __uint_as_float(u32(__float_as_uint(m.nbeta))), u32(tl.block));
    }
}

// ---------------------------------------------------------------------------
// Coupling-row staging alone: the sweep's per-step copies of J[pos] (the same
// chunks of the same sites, from the same chunked layout, by the same waves
// and instructions, one barrier per chunk) with no gathers, reduction or
// accept.  With INFLIGHT = 1 each chunk's copy is awaited before the barrier
// that follows its issue, as in the sweep; with INFLIGHT = 2 the copy of the
// chunk just issued stays in flight across the barrier (a barrier without
// the vmcnt(0) of __syncthreads; the LDS contents are not used).  The first
// is the staging time of the sweep's double-buffered launch shape, the
// second that of the copies alone, overlapped.
template <int PF, int NWAVES, int INFLIGHT>
__global__ __launch_bounds__(64 * NWAVES) void k_stage_probe(const float* __restrict__ Ju,
                                                             const uint32_t* __restrict__ pos_list,
                                                             uint32_t L, uint32_t nchunk,
                                                             uint32_t nsteps) {
    constexpr int NT = 64 * NWAVES;
    constexpr int DMA_T = NT < 512 ? NT : 512;
    constexpr uint32_t chb = PF * DMA_T * 16;
    extern __shared__ f4v lds4[];
    char* ldsb = reinterpret_cast<char*>(lds4);
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const __amdgpu_buffer_rsrc_t rJ = make_rsrc(Ju, L * nchunk * chb);
    const uint32_t voff = tid * 16;
    uint32_t buf = 0;
    for (uint32_t t = 0; t < nsteps; t++) {
        const uint32_t pos = __builtin_amdgcn_readfirstlane(pos_list[t]);
        for (uint32_t k = 0; k < nchunk; k++) {
            if (wave < DMA_T / 64)
                chunk_dma<PF, DMA_T>(rJ, voff, (pos * nchunk + k) * chb, ldsb + buf * chb, wave);
            if constexpr (INFLIGHT == 1) {
                wait_vmem();
                __syncthreads();
            } else {
                wait_vmem_upto<PF>();
                barrier_copies_in_flight();
            }
            buf ^= 1;
        }
    }
    wait_vmem();
}

// ---------------------------------------------------------------------------
// Batch energies E(s) = sum_{i<m} J_{i,m}(s_i, s_m) of nseq sequences.
// Restates getEnergies (mcmc.cl:249-345); same layout and staging as the
// sweep, streaming for each row i only the chunks holding m > i.
template <int KW, int S, int NWAVES, int PF, bool OVL>
__global__ __launch_bounds__(64 * NWAVES) void k_energies(
    const float* __restrict__ Ju, const uint8_t* __restrict__ seq, float* __restrict__ E,
    uint32_t ldseq, uint32_t nseq, uint32_t L, uint32_t q, uint32_t nchunk) {
    constexpr int NT = 64 * NWAVES;
    constexpr int K = 4 * KW;
    extern __shared__ f4v lds4[];

    // the padded chunk is PF float4 per thread of the first DMA_T threads
    // (the chunk layout's unit); waves beyond them take no part in staging
    constexpr int DMA_T = NT < 512 ? NT : 512;
    constexpr uint32_t chf = PF * DMA_T * 4;  // padded chunk (floats)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t j = lane & 31;
    const bool upper = lane >= 32;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t wfirst = blockIdx.x * (2 * S * NWAVES) + (wave * 2 + (upper ? 1 : 0)) * S;

    uint32_t st[S][KW];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const bool valid = wfirst + s < nseq;
#pragma unroll
        for (int kw = 0; kw < KW; kw++) {
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                uint32_t m = j + CH * (4 * kw + e);
                uint32_t v = (valid && m < L) ? seq[(size_t)m * ldseq + wfirst + s] : 0u;
                word |= v << (8 * e);
            }
            st[s][kw] = word;
        }
        asm volatile("" ::: "memory");
    }
    float acc[S];
#pragma unroll
    for (int s = 0; s < S; s++) acc[s] = 0.f;
    // row offsets of the slots: in LDS bytes, so that a gather address is
    // one SDWA shift of the partner residue and one add, for up to 16 slots
    // per half-wave (3.05 against 4.67 ms at L=300 q=8, 7.9 against 10.2 ms
    // at L=400 q=21); in floats for 32 slots, where byte offsets took 3.4
    // against 2.5 ms at L=232 q=21 (no launch shape uses 32 slots now)
    constexpr bool BYTE_ADDR = S <= 16;
    uint32_t base[S];

    const __amdgpu_buffer_rsrc_t rJ = make_rsrc(Ju, L * nchunk * chf * 4);
    const uint32_t voff = tid * 16;
    // staging waves, as in the sweep
    const bool stage = (PF == 1) ? (uint32_t)wave < (q * q + 7) / 8 : wave < DMA_T / 64;
    char* ldsb = reinterpret_cast<char*>(lds4);
    if (L >= 2) {
        // OVL: copy cursor (row ci, chunk ck) two units ahead of the gathers
        // and the byte offset of the unit one ahead; at the last unit the
        // cursor stays, so copies past the end repeat it into unread slots
        uint32_t ci = 0, ck = 0, off_next = 0;
        auto advance = [&]() {
            const bool wrap = ck + 1 >= nchunk;
            const bool stay = wrap && ci + 2 >= L;
            ck = stay ? ck : (wrap ? (ci + 2) >> 5 : ck + 1);
            ci = (wrap && !stay) ? ci + 1 : ci;
        };
        if constexpr (OVL) {
            static_assert(PF == 7 && DMA_T == 512, "overlapping ring: chunks of q = 17..21");
            // unit (0, 0) whole into slot 0, then the early part of the next
            if (wave < 8) {
                OverlapRing::copy<OverlapRing::EARLY_OPS>(rJ, voff, 0, ldsb, 0,
                                                          OverlapRing::early_first(0), wave);
                OverlapRing::copy<OverlapRing::LATE_OPS>(rJ, voff, 0, ldsb, 0,
                                                         OverlapRing::late_first(0), wave);
            }
            advance();
            off_next = (ci * nchunk + ck) * chf * 4;
            if (wave < 8)
                OverlapRing::copy<OverlapRing::EARLY_OPS>(rJ, voff, off_next, ldsb, 1,
                                                          OverlapRing::early_first(1), wave);
            advance();
            wait_vmem_upto<OverlapRing::EARLY_OPS>();
            barrier_copies_in_flight();
        } else {
            // first unit: row 0, chunk 0
            if (stage) chunk_dma<PF, DMA_T>(rJ, voff, 0, ldsb, wave);
            wait_vmem();
            __syncthreads();
        }
        uint32_t buf = 0;
        const uint32_t qch = q * CH;
        for (uint32_t i = 0; i + 1 < L; i++) {
            const uint32_t kp = i >> 5, owner = i & 31;
            const uint32_t kwp = kp >> 2, shp = (kp & 3) * 8;
            const uint32_t c0 = (i + 1) >> 5;
            // each slot's residue at site i, from the owner lane of its
            // half-wave (one ds_bpermute per slot; two v_readlane and a
            // select per slot took 9-19% longer), then its row offset
            const int own_addr = (int)((owner + (lane & 32)) * 4);
#pragma unroll
            for (int s = 0; s < S; s++) {
                const uint32_t av = __builtin_amdgcn_ubfe(select_word<KW>(st[s], kwp), shp, 8);
                base[s] = (uint32_t)__builtin_amdgcn_ds_bpermute(own_addr, (int)av);
            }
#pragma unroll
            for (int s = 0; s < S; s++)
                base[s] = BYTE_ADDR ? mad_u24(base[s], qch * 4, j * 4) : mad_u24(base[s], qch, j);
#pragma unroll
            for (int k = 0; k < K; k++) {
                if ((uint32_t)k >= c0 && (uint32_t)k < nchunk) {
                    const char* cbb;
                    if constexpr (OVL) {
                        // late part of the next unit, early part of the one after
                        if (wave < 8) {
                            OverlapRing::copy<OverlapRing::LATE_OPS>(
                                rJ, voff, off_next, ldsb, (buf + 1) & 3,
                                OverlapRing::late_first((buf + 1) & 3), wave);
                            OverlapRing::copy<OverlapRing::EARLY_OPS>(
                                rJ, voff, (ci * nchunk + ck) * chf * 4, ldsb, (buf + 2) & 3,
                                OverlapRing::early_first((buf + 2) & 3), wave);
                        }
                        off_next = (ci * nchunk + ck) * chf * 4;
                        advance();
                        cbb = ldsb + OverlapRing::base(buf);
                    } else {
                        const bool last_chunk = (uint32_t)k + 1 == nchunk;
                        const bool has_next = !last_chunk || (i + 2 < L);
                        if (has_next) {
                            uint32_t ni = last_chunk ? i + 1 : i;
                            uint32_t nc = last_chunk ? ((i + 2) >> 5) : (uint32_t)k + 1;
                            if (stage)
                                chunk_dma<PF, DMA_T>(rJ, voff, (ni * nchunk + nc) * chf * 4,
                                                     ldsb + (buf ^ 1) * chf * 4, wave);
                        }
                        cbb = ldsb + buf * chf * 4;
                    }
                    const float* cb = reinterpret_cast<const float*>(cbb);
                    if constexpr (S > 16) {
                        // the state is loop-invariant: without this the
                        // compiler hoists the S*K extracted residues out of
                        // the loops, which at 32 slots spills registers
#pragma unroll
                        for (int s = 0; s < S; s++)
#pragma unroll
                            for (int kw = 0; kw < KW; kw++) asm volatile("" : "+v"(st[s][kw]));
                    }
                    // on the diagonal chunk only partners m > i count: a 0/1
                    // factor (fma(v, 1, acc) == acc + v exactly)
                    const float keep = ((uint32_t)k != kp || j > owner) ? 1.f : 0.f;
                    constexpr int G = (S < 16) ? S : 16;
#pragma unroll
                    for (int s0 = 0; s0 < S; s0 += G) {
                        float v[G];
#pragma unroll
                        for (int g = 0; g < G; g++) {
                            const uint32_t w = state_word<KW>(st[s0 + g], (uint32_t)k >> 2);
                            if constexpr (BYTE_ADDR) {
                                v[g] = *reinterpret_cast<const float*>(
                                    cbb + base[s0 + g] + byte_x128(w, k & 3));
                            } else {
                                const uint32_t sm = (w >> ((k & 3) * 8)) & 0xffu;
                                v[g] = cb[base[s0 + g] + sm * CH];
                            }
                        }
#pragma unroll
                        for (int g = 0; g < G; g++) acc[s0 + g] = __fmaf_rn(v[g], keep, acc[s0 + g]);
                    }
                    if constexpr (OVL) {
                        // the next unit has landed (only the early part of
                        // the one after may be in flight)
                        wait_vmem_upto<OverlapRing::EARLY_OPS>();
                        barrier_copies_in_flight();
                        buf = (buf + 1) & 3;
                    } else {
                        wait_vmem();
                        __syncthreads();
                        buf ^= 1;
                    }
                }
            }
        }
    }
    // copies issued past the last unit
    if constexpr (OVL) wait_vmem();
#pragma unroll
    for (int s = 0; s < S; s++) {
        float e = half_wave_sum(acc[s]);
        if (j == 0 && wfirst + s < nseq) E[wfirst + s] = e;
    }
}

}  // namespace mi3
