// Shared device helpers for the MI355X (gfx950) Mi3 engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi3 {

// Random-stream tags of the 128-bit Philox counter (word 3).
constexpr uint32_t TAG_WALK = 0x57414C4Bu;   // per-walker proposal / acceptance draws
constexpr uint32_t TAG_POS = 0x504F5349u;    // shared site sequence
constexpr uint32_t TAG_INDEP = 0x494E4450u;  // independent-site reseed
constexpr uint32_t TAG_MADE = 0x4D414445u;   // autoregressive proposal draws
constexpr uint32_t TAG_MHAC = 0x4D484143u;   // independence Metropolis-Hastings accepts

// Sites per LDS chunk and lanes per walker group in the sweep/energy kernels.
constexpr int CH = 32;

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter-based: the
// output depends only on (counter, key), so every walker owns an independent
// stream addressed by (step, walker) with no state kept between launches.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Uniform integer in [0, n): floor(x * n / 2^32).
__device__ __forceinline__ uint32_t mulhi_range(uint32_t x, uint32_t n) { return __umulhi(x, n); }

// Packed pair index of (i, j), i < j, in the (L(L-1)/2, q*q) coupling format.
__device__ __forceinline__ uint32_t pair_index(uint32_t i, uint32_t j, uint32_t L) {
    return i * (2 * L - i - 1) / 2 + (j - i - 1);
}

// 32-lane butterfly exchange (ds_swizzle bitmask mode, xor within each
// half-wave).  Pattern bits: and_mask = 0x1f, or_mask = 0, xor_mask = M.
template <int M>
__device__ __forceinline__ float swz_xor(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (M << 10) | 0x1f));
}

__device__ __forceinline__ float half_wave_sum(float v) {
    v += swz_xor<16>(v);
    v += swz_xor<8>(v);
    v += swz_xor<4>(v);
    v += swz_xor<2>(v);
    v += swz_xor<1>(v);
    return v;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

}  // namespace mi3
