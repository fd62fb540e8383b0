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
//   * chunks are double-buffered: the next chunk is loaded into registers
//     while the current one is gathered, one barrier per chunk.
//
// Chunked coupling layout (built by k_unpack_chunked in k_layout.hip):
//   Ju[i][c][a*q + b][jl] = J_{i,m}(a, b),  m = 32c + jl,
// symmetric (both i<m and i>m filled), zero for m == i and m >= L.
#include "common.h"

namespace mi3 {

// Native 4-float vector (HIP's float4 is a union-based struct, which keeps
// arrays of it out of registers).
typedef float f4v __attribute__((ext_vector_type(4)));

// Chunks are padded to PF * NT float4, so staging is unconditional.
template <int PF, int NT>
__device__ __forceinline__ void chunk_load(f4v (&pf)[PF], const f4v* __restrict__ src, int tid) {
#pragma unroll
    for (int r = 0; r < PF; r++) pf[r] = src[tid + r * NT];
}

template <int PF, int NT>
__device__ __forceinline__ void chunk_store(const f4v (&pf)[PF], f4v* dst, int tid) {
#pragma unroll
    for (int r = 0; r < PF; r++) dst[tid + r * NT] = pf[r];
}

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
template <int KW, int S, int NWAVES, int PF>
__global__ __launch_bounds__(64 * NWAVES) void k_sweep(
    const float* __restrict__ Ju, uint8_t* __restrict__ seq, const uint32_t* __restrict__ pos_list,
    uint32_t ldseq, uint32_t L, uint32_t q, uint32_t nchunk, uint32_t nsteps, uint32_t k0,
    uint32_t k1, uint64_t step0, uint32_t walker_base, float nbeta) {
    constexpr int NT = 64 * NWAVES;
    constexpr int K = 4 * KW;
    extern __shared__ f4v lds4[];
    float* lds = reinterpret_cast<float*>(lds4);

    constexpr uint32_t chf = PF * NT * 4;  // padded chunk (floats)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t j = lane & 31;
    const bool upper = lane >= 32;
    const int wave = tid >> 6;
    const uint32_t wfirst = blockIdx.x * (2 * S * NWAVES) + (wave * 2 + (upper ? 1 : 0)) * S;

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

    uint32_t P[S];      // packed (u24 << 8 | b) for 32 steps, lane (t>>1) + 16*(t&1)
    uint32_t cur[S];    // packed draw of the current step
    uint32_t base[S];   // (a*q*32 + j) | (b*q*32 + j) << 16, LDS float index
    float acc[S];

    f4v pf[PF];
    // prologue: chunk 0 of step 0 into buffer 0
    if (nsteps == 0) return;
    chunk_load<PF, NT>(pf, reinterpret_cast<const f4v*>(Ju + (size_t)pos_list[0] * nchunk * chf),
                       tid);
    chunk_store<PF, NT>(pf, lds4, tid);
    __syncthreads();
    uint32_t buf = 0;
    const uint32_t qch = q * CH;

    for (uint32_t t = 0; t < nsteps; t++) {
        const uint32_t pos = pos_list[t];
        const uint32_t kp = pos >> 5, owner = pos & 31;
        const uint32_t kwp = kp >> 2, shp = (kp & 3) * 8;
        const uint32_t tl = t & 31;
        if (tl == 0) {
            // refill the per-walker draws for steps t .. t+31 (global step even)
            const uint64_t h = ((step0 + t) >> 1) + (j & 15);
#pragma unroll
            for (int s = 0; s < S; s++) {
                uint4 r = philox4x32_10(make_uint4((uint32_t)h, (uint32_t)(h >> 32),
                                                   walker_base + wfirst + s, TAG_WALK),
                                        k0, k1);
                uint32_t x = (j < 16) ? r.x : r.z;
                uint32_t y = (j < 16) ? r.y : r.w;
                P[s] = ((y >> 8) << 8) | mulhi_range(x, q);
            }
        }
        const uint32_t rl = (tl >> 1) + ((tl & 1) << 4);
#pragma unroll
        for (int s = 0; s < S; s++) {
            uint32_t av = (select_word<KW>(st[s], kwp) >> shp) & 0xffu;
            uint32_t aA = __builtin_amdgcn_readlane(av, owner);
            uint32_t aB = __builtin_amdgcn_readlane(av, owner + 32);
            uint32_t a = upper ? aB : aA;
            uint32_t pA = __builtin_amdgcn_readlane(P[s], rl);
            uint32_t pB = __builtin_amdgcn_readlane(P[s], rl + 32);
            cur[s] = upper ? pB : pA;
            uint32_t b = cur[s] & 0xffu;
            base[s] = (a * qch + j) | ((b * qch + j) << 16);
            acc[s] = 0.f;
        }
        // stream the couplings of site pos, chunk by chunk
#pragma unroll
        for (int k = 0; k < K; k++) {
            if ((uint32_t)k < nchunk) {
                const bool last_chunk = (uint32_t)k + 1 == nchunk;
                const bool has_next = !last_chunk || (t + 1 < nsteps);
                if (has_next) {
                    uint32_t npos = last_chunk ? pos_list[t + 1] : pos;
                    uint32_t nc = last_chunk ? 0u : (uint32_t)k + 1;
                    chunk_load<PF, NT>(
                        pf, reinterpret_cast<const f4v*>(Ju + ((size_t)npos * nchunk + nc) * chf),
                        tid);
                }
                const float* cb = lds + buf * chf;
#pragma unroll
                for (int s = 0; s < S; s++) {
                    uint32_t sm = (select_word<KW>(st[s], (uint32_t)k >> 2) >> ((k & 3) * 8)) & 0xffu;
                    uint32_t off = sm * CH;
                    float vA = cb[(base[s] & 0xffffu) + off];
                    float vB = cb[(base[s] >> 16) + off];
                    acc[s] += vB - vA;
                }
                if (has_next) chunk_store<PF, NT>(pf, lds4 + (buf ^ 1) * (chf / 4), tid);
                __syncthreads();
                buf ^= 1;
            }
        }
        // accept / reject
#pragma unroll
        for (int s = 0; s < S; s++) {
            float dE = half_wave_sum(acc[s]);
            float u = (float)(cur[s] >> 8) * 0x1.0p-24f;
            const bool upd = (__expf(nbeta * dE) > u) && (j == owner);
            const uint32_t b = cur[s] & 0xffu;
#pragma unroll
            for (int kw = 0; kw < KW; kw++) {
                uint32_t nw = (st[s][kw] & ~(0xffu << shp)) | (b << shp);
                asm volatile("" : "+v"(nw));
                st[s][kw] = (upd && kwp == (uint32_t)kw) ? nw : st[s][kw];
            }
        }
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

// ---------------------------------------------------------------------------
// Batch energies E(s) = sum_{i<m} J_{i,m}(s_i, s_m) of nseq sequences.
// Restates getEnergies (mcmc.cl:249-345); same layout and staging as the
// sweep, streaming for each row i only the chunks holding m > i.
template <int KW, int S, int NWAVES, int PF>
__global__ __launch_bounds__(64 * NWAVES) void k_energies(
    const float* __restrict__ Ju, const uint8_t* __restrict__ seq, float* __restrict__ E,
    uint32_t ldseq, uint32_t nseq, uint32_t L, uint32_t q, uint32_t nchunk) {
    constexpr int NT = 64 * NWAVES;
    constexpr int K = 4 * KW;
    extern __shared__ f4v lds4[];
    float* lds = reinterpret_cast<float*>(lds4);

    constexpr uint32_t chf = PF * NT * 4;  // padded chunk (floats)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t j = lane & 31;
    const bool upper = lane >= 32;
    const int wave = tid >> 6;
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
    uint32_t base[S];

    f4v pf[PF];
    if (L >= 2) {
        // first unit: row 0, chunk 0
        chunk_load<PF, NT>(pf, reinterpret_cast<const f4v*>(Ju), tid);
        chunk_store<PF, NT>(pf, lds4, tid);
        __syncthreads();
        uint32_t buf = 0;
        const uint32_t qch = q * CH;
        for (uint32_t i = 0; i + 1 < L; i++) {
            const uint32_t kp = i >> 5, owner = i & 31;
            const uint32_t kwp = kp >> 2, shp = (kp & 3) * 8;
            const uint32_t c0 = (i + 1) >> 5;
#pragma unroll
            for (int s = 0; s < S; s++) {
                uint32_t av = (select_word<KW>(st[s], kwp) >> shp) & 0xffu;
                uint32_t aA = __builtin_amdgcn_readlane(av, owner);
                uint32_t aB = __builtin_amdgcn_readlane(av, owner + 32);
                base[s] = (upper ? aB : aA) * qch + j;
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                if ((uint32_t)k >= c0 && (uint32_t)k < nchunk) {
                    const bool last_chunk = (uint32_t)k + 1 == nchunk;
                    const bool has_next = !last_chunk || (i + 2 < L);
                    if (has_next) {
                        uint32_t ni = last_chunk ? i + 1 : i;
                        uint32_t nc = last_chunk ? ((i + 2) >> 5) : (uint32_t)k + 1;
                        chunk_load<PF, NT>(
                            pf, reinterpret_cast<const f4v*>(Ju + ((size_t)ni * nchunk + nc) * chf),
                            tid);
                    }
                    const float* cb = lds + buf * chf;
                    const bool diag = (uint32_t)k == kp;
#pragma unroll
                    for (int s = 0; s < S; s++) {
                        uint32_t sm = (select_word<KW>(st[s], (uint32_t)k >> 2) >> ((k & 3) * 8)) & 0xffu;
                        float v = cb[base[s] + sm * CH];
                        if (diag) v = (j > owner) ? v : 0.f;
                        acc[s] += v;
                    }
                    if (has_next) chunk_store<PF, NT>(pf, lds4 + (buf ^ 1) * (chf / 4), tid);
                    __syncthreads();
                    buf ^= 1;
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        float e = half_wave_sum(acc[s]);
        if (j == 0 && wfirst + s < nseq) E[wfirst + s] = e;
    }
}

}  // namespace mi3
