// Layout, fill, random-site and independent-reseed kernels.
#include "common.h"

namespace mi3 {

// Packed (L(L-1)/2, q*q) couplings -> chunked symmetric layout
// Ju[i][c][a*q+b][jl] = J_{i,m}(a,b), m = 32c + jl (zero for m == i, m >= L).
// Same role as unpackfV (mcmc.cl:43-65): the (j,i) block is the transpose of
// the stored (i,j) block; the reference leaves the diagonal unfilled and
// skips it, here it is zero so no skip is needed.
// Every chunk is padded to `chf` floats (zeros after the q*q*32 used ones).
__global__ void k_unpack_chunked(const float* __restrict__ J, float* __restrict__ Ju, uint32_t L,
                                 uint32_t q, uint32_t nchunk, uint32_t chf, uint64_t total) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint32_t qq = q * q;
    uint64_t unit = idx / chf;
    uint32_t within = (uint32_t)(idx - unit * chf);
    uint32_t c = (uint32_t)(unit % nchunk);
    uint32_t i = (uint32_t)(unit / nchunk);
    uint32_t jl = within & 31;
    uint32_t ac = within >> 5;
    uint32_t m = c * CH + jl;
    float v = 0.f;
    if (ac < qq && m < L && m != i) {
        uint32_t a = ac / q, b = ac - a * q;
        if (i < m)
            v = J[(size_t)pair_index(i, m, L) * qq + a * q + b];
        else
            v = J[(size_t)pair_index(m, i, L) * qq + b * q + a];
    }
    Ju[idx] = v;
}

// Plain streaming kernels over n 16-byte words (grid-stride, eight
// nontemporal loads in flight per thread): the references for the HBM
// bandwidth a kernel reaches.  The read kernel stores the xor of the data
// only if it equals a key the caller passes, so that the loads stay.
typedef uint32_t u4v __attribute__((ext_vector_type(4)));
constexpr int STREAM_U = 8;

__global__ __launch_bounds__(256) void k_stream_copy(const u4v* __restrict__ src,
                                                     u4v* __restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (STREAM_U - 1) * stride < n; i += STREAM_U * stride) {
        u4v v[STREAM_U];
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) v[u] = __builtin_nontemporal_load(&src[i + u * stride]);
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) __builtin_nontemporal_store(v[u], &dst[i + u * stride]);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_stream_read(const u4v* __restrict__ src, uint64_t n,
                                                     uint32_t key, uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (; i + (STREAM_U - 1) * stride < n; i += STREAM_U * stride) {
        u4v v[STREAM_U];
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) v[u] = __builtin_nontemporal_load(&src[i + u * stride]);
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    for (; i < n; i += stride) {
        const u4v a = src[i];
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    if (acc == key) out[0] = acc;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t val, uint64_t n) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) p[idx] = val;
}

// Shared site sequence for global steps step0 .. step0+n-1:
// pos = mulhi(Philox(key, (g, 0, TAG_POS)).x, L).  Replaces the host
// RandomState position list (mcmcGPU.py:447-458).
__global__ void k_positions(uint32_t* __restrict__ pos, uint32_t n, uint32_t L, uint32_t k0,
                            uint32_t k1, uint64_t step0) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint64_t g = step0 + t;
    uint4 r = philox4x32_10(make_uint4((uint32_t)g, (uint32_t)(g >> 32), 0u, TAG_POS), k0, k1);
    pos[t] = mulhi_range(r.x, L);
}

// Independent-site reseed (gen_indep, mcmc.cl:170-244): residue at site i is
// the number of cumulative probabilities cprob[i][0..q-2] that the uniform
// u = (r >> 8) * 2^-24 reaches, which is what the reference's binary search
// returns for a monotone row.  One thread per walker, site-major output.
__global__ void k_gen_indep(uint8_t* __restrict__ seq, uint32_t ldseq, uint32_t nwalk,
                            const float* __restrict__ cprob, uint32_t L, uint32_t q, uint32_t k0,
                            uint32_t k1, uint32_t call, uint32_t walker_base) {
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwalk) return;
    for (uint32_t k = 0; 4 * k < L; k++) {
        uint4 r = philox4x32_10(make_uint4(k, call, walker_base + w, TAG_INDEP), k0, k1);
        uint32_t rr[4] = {r.x, r.y, r.z, r.w};
        for (uint32_t e = 0; e < 4 && 4 * k + e < L; e++) {
            uint32_t i = 4 * k + e;
            float u = (float)(rr[e] >> 8) * 0x1.0p-24f;
            uint32_t res = 0;
            const float* cp = cprob + (size_t)i * (q - 1);
            for (uint32_t a = 0; a + 1 < q; a++) res += (u >= cp[a]) ? 1u : 0u;
            seq[(size_t)i * ldseq + w] = (uint8_t)res;
        }
    }
}

// Fill every walker with the same sequence (fillSeqs, mcmcGPU.py:927-932).
__global__ void k_fill_seqs(uint8_t* __restrict__ seq, uint32_t ldseq, uint32_t nwalk,
                            const uint8_t* __restrict__ s, uint32_t L) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)L * nwalk) return;
    uint32_t i = (uint32_t)(idx / nwalk), w = (uint32_t)(idx - (uint64_t)i * nwalk);
    seq[(size_t)i * ldseq + w] = s[i];
}

// Subsequence substitution (copySubseq, mcmc.cl:98-122): every sequence of
// dst takes, at each marked site, the residue of sequence `seqind` of src.
__global__ void k_copy_subseq(uint8_t* __restrict__ dst, uint32_t lddst, uint32_t ndst,
                              const uint8_t* __restrict__ src, uint32_t ldsrc, uint32_t seqind,
                              const uint8_t* __restrict__ marks, uint32_t L) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)L * ndst) return;
    uint32_t i = (uint32_t)(idx / ndst), w = (uint32_t)(idx - (uint64_t)i * ndst);
    if (marks[i]) dst[(size_t)i * lddst + w] = src[(size_t)i * ldsrc + seqind];
}

}  // namespace mi3
