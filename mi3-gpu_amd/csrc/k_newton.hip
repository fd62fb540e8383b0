// Quasi-Newton (Zwanzig reweighting) kernels of one inference round.
#include "common.h"

#include <utility>

namespace mi3 {

// bimarg = float(count) / float(nseq), correctly rounded (mcmc.cl:518-528).
__global__ void k_bicounts_to_bimarg(const uint32_t* __restrict__ c, float* __restrict__ bi,
                                     uint32_t nseq, uint64_t n) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) bi[idx] = __fdiv_rn((float)c[idx], (float)nseq);
}

// Per q x q block: coupling update (updatedJ, mcmc.cl:829-843)
//   dJ <- dJ - gamma*(f_target - f)/(f + pc)
// optionally followed by the zero-gauge L1/L2 step (reg_l1z / reg_l2z,
// mcmc.cl:845-902, with zeroGauge 764-822):
//   J0 = zerogauge(J + dJ); R = -lam*sign(J0)*gamma/(f+pc)   (l1)
//                            R = -lam*J0*gamma/(f+pc)        (l2)
//   dJ <- (sign(J0) != sign(J0+R)) ? dJ - J0 : dJ + R
// One wave per pair block; `mode` 0 = update only, 1 = reg only (l1),
// 2 = reg only (l2), 3 = update + l1, 4 = update + l2.
__device__ __forceinline__ float fsign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

__global__ __launch_bounds__(256) void k_update_reg(const float* __restrict__ bt,
                                                    const float* __restrict__ bi,
                                                    const float* __restrict__ J,
                                                    float* __restrict__ dJ, uint32_t nP,
                                                    uint32_t q, float gamma, float pc, float lam,
                                                    int mode) {
    __shared__ float sJ[4][32 * 32];
    __shared__ float sh[4][2][32];
    const uint32_t wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * 4 + wave;
    const uint32_t qq = q * q;
    const bool active = p < nP;
    const bool upd = (mode == 0) || (mode >= 3);
    const int reg = (mode == 1 || mode == 3) ? 1 : ((mode == 2 || mode == 4) ? 2 : 0);
    float dj[16];  // q*q <= 1024 -> at most 16 per lane
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint32_t e = l + 64 * r;
        dj[r] = 0.f;
        if (active && e < qq) {
            size_t o = (size_t)p * qq + e;
            float d = dJ[o];
            if (upd) {
                float f = bi[o];
                d = __fsub_rn(d, __fdiv_rn(__fmul_rn(gamma, __fsub_rn(bt[o], f)), __fadd_rn(f, pc)));
            }
            dj[r] = d;
            if (reg) sJ[wave][e] = __fadd_rn(J[o], d);
        }
    }
    if (reg) {
        __syncthreads();
        // row means (lanes 0..q-1) and column means (lanes 32..32+q-1)
        if (active) {
            if (l < q) {
                float s = 0.f;
                for (uint32_t b = 0; b < q; b++) s += sJ[wave][l * q + b];
                sh[wave][0][l] = s / q;
            } else if (l >= 32 && l - 32 < q) {
                uint32_t b = l - 32;
                float s = 0.f;
                for (uint32_t a = 0; a < q; a++) s += sJ[wave][a * q + b];
                sh[wave][1][b] = s / q;
            }
        }
        __syncthreads();
        if (active) {
            float mean = 0.f;
            for (uint32_t a = 0; a < q; a++) mean += sh[wave][0][a];
            mean /= q;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                uint32_t e = l + 64 * r;
                if (e < qq) {
                    uint32_t a = e / q, b = e - a * q;
                    float hi = sh[wave][0][a] - mean, hj = sh[wave][1][b] - mean;
                    float J0 = sJ[wave][e] - (hi + hj) - mean;
                    size_t o = (size_t)p * qq + e;
                    float R = (reg == 1) ? -lam * fsign(J0) * gamma / (bi[o] + pc)
                                         : -lam * J0 * gamma / (bi[o] + pc);
                    dj[r] = (fsign(J0) != fsign(J0 + R)) ? dj[r] - J0 : dj[r] + R;
                }
            }
        }
    }
    if (active) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            uint32_t e = l + 64 * r;
            if (e < qq) dJ[(size_t)p * qq + e] = dj[r];
        }
    }
}

// Divide every pair block by its sum (renormalize_bimarg, mcmc.cl:729-759).
__global__ __launch_bounds__(256) void k_renormalize(float* __restrict__ bi, uint32_t nP,
                                                     uint32_t qq) {
    const uint32_t wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * 4 + wave;
    if (p >= nP) return;
    float* b = bi + (size_t)p * qq;
    float s = 0.f;
    for (uint32_t e = l; e < qq; e += 64) s += b[e];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    for (uint32_t e = l; e < qq; e += 64) b[e] = b[e] / s;
}

// Two-pass deterministic reductions over n floats.  Pass 1 writes one
// partial per block: min (starting from 0, as minFloats, mcmc.cl:558-584),
// sum w, sum w^2 (weight_stats, mcmc.cl:586-622) in float64.
__global__ __launch_bounds__(256) void k_reduce_partial(const float* __restrict__ x, uint32_t n,
                                                        int op, double* __restrict__ part) {
    __shared__ double s1[256], s2[256];
    uint32_t tid = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (uint32_t i = blockIdx.x * 256 + tid; i < n; i += gridDim.x * 256) {
        double v = x[i];
        if (op == 0)
            a = fmin(a, v);
        else {
            a += v;
            b += v * v;
        }
    }
    s1[tid] = a;
    s2[tid] = b;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            if (op == 0)
                s1[tid] = fmin(s1[tid], s1[tid + off]);
            else {
                s1[tid] += s1[tid + off];
                s2[tid] += s2[tid + off];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ part,
                                                      uint32_t nb, int op,
                                                      float* __restrict__ out) {
    __shared__ double s1[256], s2[256];
    uint32_t tid = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (uint32_t i = tid; i < nb; i += 256) {
        if (op == 0)
            a = fmin(a, part[2 * i]);
        else {
            a += part[2 * i];
            b += part[2 * i + 1];
        }
    }
    s1[tid] = a;
    s2[tid] = b;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            if (op == 0)
                s1[tid] = fmin(s1[tid], s1[tid + off]);
            else {
                s1[tid] += s1[tid + off];
                s2[tid] += s2[tid + off];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = (float)s1[0];
        if (op != 0) out[1] = (float)s2[0];
    }
}

// w = exp(-dE + ref) (dE_to_weights, mcmc.cl:624-633).
__global__ void k_dE_to_weights(const float* __restrict__ dE, float* __restrict__ w, float ref,
                                uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = expf(-dE[i] + ref);
}

// w = exp((beta-1)*(E - ref)) (fixed_beta_weights, mcmc.cl:635-644).
__global__ void k_fixed_beta_weights(const float* __restrict__ E, float* __restrict__ w,
                                     float ref, float bm1, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = expf(bm1 * (E[i] - ref));
}

// dst += src (addFloatBufs, mcmc.cl:720-727).
__global__ void k_add_floats(float* __restrict__ dst, const float* __restrict__ src, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

// Weighted pair histograms F_{im}(a,b) = sum_n w_n [s_ni = a][s_nm = b]
// (weightedMarg, mcmc.cl:646-718), unnormalised.
//
// Accumulation is exact integer arithmetic in 64-bit fixed point: every
// weight becomes wq_n = round(w_n * 2^62 / sum(w)), so that no bin of any
// subset of sequences can exceed 2^63; bins are summed with LDS ds_add_u64
// (integer LDS adds run about 20x faster than float ones), parts of the
// sequence range are added with 64-bit integer atomics, and the result is
// scaled back to float once.  The rounding error per term is below
// 2^-63 sum(w) (plus the 2^-53 relative rounding of the double product),
// so every bin, even one holding a single sequence of a 2^24-sequence set,
// is the correctly rounded float of its exact weighted sum to within an ulp,
// and the result does not depend on the order of any addition.
//
// Pass 1: fixed-point weights (a non-finite or zero total makes every output
// NaN, which keeps the divergence check of the Newton loop working), in the
// sequences' order.
__global__ void k_weights_fixed(const float* __restrict__ w, uint32_t n,
                                const float* __restrict__ wsum, uint64_t* __restrict__ wq,
                                uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double tot = (double)wsum[0];
    const bool ok = tot > 0.0 && tot < INFINITY;
    if (i == 0) bad[0] = ok ? 0u : 1u;
    if (i >= n) return;
    const double v = ok ? (double)w[i] * (4611686018427387904.0 / tot) : 0.0;  // 2^62 / sum
    wq[i] = (v > 0.0) ? (uint64_t)(v + 0.5) : 0ull;
}

// Pass 2: one work-group per (unit of 32 consecutive pairs, part of the
// sequence range).  Unit u holds pairs 32u .. 32u + 31 of the packed order,
// rows chained (a unit can span rows), so only the last unit has columns
// with no pair.  The 32 pair histograms live in LDS as uint64 [a*q+b][32]:
// lane l of a wave owns pair 32u + (l & 31).  A wave takes groups of 32
// sequences, the lower half-wave the first 16 and the upper half-wave the
// next 16, so that one 16-byte load per lane of each half-wave reads both
// halves of a 32-byte piece of the same row: each cache line looked up
// serves 16 adds per lane.  (With two half-waves adding the even and odd
// sequences of one group of 16, each lookup served 8, and the vector memory
// path, not the LDS atomics, set the pace.)  The 16 waves take groups in
// turn, two groups per iteration with all their loads issued before the
// first add.  The part's histograms are added into the uint64 accumulator
// (zeroed beforehand) with integer atomics; the accumulator holds the pairs
// from pair_lo on (a run of units covers a contiguous pair range).
constexpr uint32_t WM_THREADS = 1024;
constexpr uint32_t WM_COLS = 32;

// One group for one lane: the residues of its half-wave's 16 sequences at
// the pair's partner site and row site, and weight (lane % 16) of those
// sequences; the add of sequence e takes weight e from lane e of its 16-lane
// row by DPP (row_newbcast), so that a lane loads one weight, not 16.
struct WmGroup {
    uint4 sm;     // partner-site residues of the half-wave's 16 sequences
    uint4 si;     // row-site residues of the same sequences
    uint64_t w;   // weight (lane % 16) of those sequences
};

// half: 16 for the upper half-wave, else 0; wsel = half + lane % 16
__device__ __forceinline__ void wm_load(WmGroup& G, const uint8_t* rowi, const uint8_t* rowm,
                                        const uint64_t* wq, uint32_t g, uint32_t half,
                                        uint32_t wsel) {
    G.sm = *reinterpret_cast<const uint4*>(rowm + g * 32 + half);
    G.si = *reinterpret_cast<const uint4*>(rowi + g * 32 + half);
    G.w = wq[(size_t)g * 32 + wsel];
}

// v from lane e of each row of 16 lanes (DPP row_newbcast:e; every source
// lane is active, so the destination needs no initial value)
#define MI3_BC(E)                                                                       \
    case E:                                                                             \
        lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)v, 0x150 + E, 0xF, 0xF, false); \
        hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), 0x150 + E, 0xF, 0xF, \
                                                false);                                 \
        break;
__device__ __forceinline__ uint64_t row_bcast64(uint64_t v, int e) {
    uint32_t lo, hi;
    switch (e) {
        MI3_BC(0) MI3_BC(1) MI3_BC(2) MI3_BC(3) MI3_BC(4) MI3_BC(5) MI3_BC(6) MI3_BC(7)
        MI3_BC(8) MI3_BC(9) MI3_BC(10) MI3_BC(11) MI3_BC(12) MI3_BC(13) MI3_BC(14)
        default: MI3_BC(15)
    }
    return ((uint64_t)hi << 32) | lo;
}
#undef MI3_BC

// n * byte B of w in one VALU op (an SDWA byte select on the multiplier),
// and 256 * byte B of w as a shift with the same select
#define MI3_SDWA_BYTE(B)                                                                     \
    template <>                                                                              \
    __device__ __forceinline__ uint32_t mul_byte<B>(uint32_t n, uint32_t w) {                \
        uint32_t r;                                                                          \
        asm("v_mul_u32_u24_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD "             \
            "src0_sel:DWORD src1_sel:BYTE_" #B : "=v"(r) : "v"(n), "v"(w));                  \
        return r;                                                                            \
    }                                                                                        \
    template <>                                                                              \
    __device__ __forceinline__ uint32_t byte_x256<B>(uint32_t w) {                           \
        uint32_t r;                                                                          \
        asm("v_lshlrev_b32_sdwa %0, 8, %1 dst_sel:DWORD dst_unused:UNUSED_PAD "              \
            "src0_sel:DWORD src1_sel:BYTE_" #B : "=v"(r) : "v"(w));                          \
        return r;                                                                            \
    }
template <int B>
__device__ __forceinline__ uint32_t mul_byte(uint32_t n, uint32_t w);
template <int B>
__device__ __forceinline__ uint32_t byte_x256(uint32_t w);
MI3_SDWA_BYTE(0)
MI3_SDWA_BYTE(1)
MI3_SDWA_BYTE(2)
MI3_SDWA_BYTE(3)
#undef MI3_SDWA_BYTE

// hcol: the histogram's base plus this lane's column offset (formed once, so
// that an address is one add of the row and residue offsets to it)
template <int W>
__device__ __forceinline__ uint32_t word(const uint4& v) {
    return W == 0 ? v.x : W == 1 ? v.y : W == 2 ? v.z : v.w;
}
template <int E>
__device__ __forceinline__ void wm_add1(const WmGroup& G, char* hcol, uint32_t qcol8) {
    const uint32_t arow = mul_byte<E & 3>(qcol8, word<E / 4>(G.si));
    const uint32_t bcol = byte_x256<E & 3>(word<E / 4>(G.sm));
    atomicAdd(reinterpret_cast<unsigned long long*>(hcol + (arow + bcol)),
              (unsigned long long)row_bcast64(G.w, E));
}
template <int... E>
__device__ __forceinline__ void wm_add_seq(const WmGroup& G, char* hcol, uint32_t qcol8,
                                           std::integer_sequence<int, E...>) {
    (wm_add1<E>(G, hcol, qcol8), ...);
}
__device__ __forceinline__ void wm_add(const WmGroup& G, char* hcol, uint32_t qcol8) {
    wm_add_seq(G, hcol, qcol8, std::make_integer_sequence<int, 16>{});
}

__global__ __launch_bounds__(WM_THREADS) void k_weighted_marg(const uint8_t* __restrict__ seq,
                                                              uint32_t ldseq, uint32_t nseq,
                                                              const uint64_t* __restrict__ wq,
                                                              uint32_t L, uint32_t q,
                                                              const uint32_t* __restrict__ units,
                                                              uint32_t nsplit,
                                                              unsigned long long* __restrict__ acc,
                                                              uint32_t pair_lo) {
    extern __shared__ unsigned long long hist64[];
    constexpr uint32_t NW = WM_THREADS / 64;
    const uint32_t unit = units[blockIdx.x / nsplit], part = blockIdx.x % nsplit;
    const uint32_t qq = q * q;
    const uint32_t npairs = L * (L - 1) / 2;
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = tid; e < qq * WM_COLS; e += WM_THREADS) hist64[e] = 0ull;
    __syncthreads();
    const uint32_t lane = tid & 63, col = lane & (WM_COLS - 1);
    const bool upper = lane >= 32;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));
    const uint32_t p0 = unit * WM_COLS;
    const bool valid = p0 + col < npairs;
    // sites (i, m) of the lane's pair: i is the last row whose first pair
    // index i(2L - i - 1)/2 is <= p.  Columns past the last pair count the
    // last pair again and are not flushed.
    const uint32_t p = valid ? p0 + col : npairs - 1;
    uint32_t lo = 0, hi = L - 1;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (mid * (2 * L - mid - 1) / 2 <= p) lo = mid;
        else hi = mid;
    }
    const uint32_t i = lo, m = p - lo * (2 * L - lo - 1) / 2 + lo + 1;
    const uint8_t* rowi = seq + (size_t)i * ldseq;
    const uint8_t* rowm = seq + (size_t)m * ldseq;
    const uint32_t qcol8 = q * WM_COLS * 8, col8 = col * 8;
    const uint32_t half = upper ? 16u : 0u, wsel = half + (lane & 15u);
    char* hcol = reinterpret_cast<char*>(hist64) + col8;
    // this part's groups of 32 sequences
    const uint32_t ngroups = nseq / 32;
    const uint32_t per = (ngroups + nsplit - 1) / nsplit;
    const uint32_t g0 = part * per;
    const uint32_t g1 = (g0 + per < ngroups) ? g0 + per : ngroups;
    // every lane adds, also those of columns past the last pair: the weight
    // broadcasts need all lanes active, and the flush below skips those
    // columns' bins
    // (two groups per iteration, g and g + NW, the next iteration's loads
    // issued before this one's adds.  Loads past the part's last group read
    // that group again and are not added, so no load is under a branch: the
    // compiler then waits for each group's own loads only.  The group bounds
    // are wave-uniform.)
    uint32_t g = g0 + wave;
    if (g < g1) {
        auto clamp = [g1](uint32_t x) { return x < g1 ? x : g1 - 1; };
        WmGroup A, B;
        wm_load(A, rowi, rowm, wq, g, half, wsel);
        wm_load(B, rowi, rowm, wq, clamp(g + NW), half, wsel);
        for (;;) {
            const uint32_t gn = g + 2 * NW;
            WmGroup An, Bn;
            wm_load(An, rowi, rowm, wq, clamp(gn), half, wsel);
            wm_load(Bn, rowi, rowm, wq, clamp(gn + NW), half, wsel);
            wm_add(A, hcol, qcol8);
            if (g + NW < g1) wm_add(B, hcol, qcol8);
            if (gn >= g1) break;
            A = An;
            B = Bn;
            g = gn;
        }
    }
    // the last nseq % 32 sequences, by the lower half of the first wave of
    // the last part
    if (part + 1 == nsplit && wave == 0 && valid && !upper) {
        for (uint32_t n = ngroups * 32; n < nseq; n++)
            atomicAdd(&hist64[__umul24(rowi[n], q * WM_COLS) + rowm[n] * WM_COLS + col],
                      (unsigned long long)wq[n]);
    }
    __syncthreads();
    for (uint32_t e = tid; e < qq * WM_COLS; e += WM_THREADS) {
        const uint32_t c = e & (WM_COLS - 1), ab = e / WM_COLS;
        const unsigned long long v = hist64[e];
        if (p0 + c < npairs && v) atomicAdd(&acc[(size_t)(p0 + c - pair_lo) * qq + ab], v);
    }
}

// Measurement probe: the 64-bit LDS atomic rate of k_weighted_marg's add
// pattern alone.  The same work-group size and histogram (q*q rows of 32
// uint64 in LDS); lane l adds into column l % 32 of a pseudo-random row (a
// 32-bit LCG per thread), so the two half-waves of a wave share columns as
// in the kernel; no global loads.  Each thread makes `iters` adds; the
// histogram's sum goes to out[0] so that nothing is dropped.
__global__ __launch_bounds__(WM_THREADS) void k_wm_atomic_probe(uint32_t q, uint32_t iters,
                                                                unsigned long long* out) {
    extern __shared__ unsigned long long hist64[];
    const uint32_t qq = q * q, tid = threadIdx.x, col = tid & (WM_COLS - 1);
    for (uint32_t e = tid; e < qq * WM_COLS; e += WM_THREADS) hist64[e] = 0ull;
    __syncthreads();
    uint32_t r = (blockIdx.x * WM_THREADS + tid) * 2654435761u + 12345u;
    for (uint32_t it = 0; it < iters; it++) {
        r = r * 1664525u + 1013904223u;
        const uint32_t row = __umulhi(r, qq);  // in [0, q*q)
        atomicAdd(&hist64[row * WM_COLS + col], (unsigned long long)(r | 1u));
    }
    __syncthreads();
    unsigned long long sum = 0;
    for (uint32_t e = tid; e < qq * WM_COLS; e += WM_THREADS) sum += hist64[e];
    atomicAdd(out, sum);
}

// Pass 3: bins of 2^-62 sum(w) to float (NaN if the total was bad).
__global__ void k_fixed_to_float(const unsigned long long* __restrict__ acc,
                                 float* __restrict__ out, uint64_t n,
                                 const float* __restrict__ wsum,
                                 const uint32_t* __restrict__ bad) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const double back = bad[0] ? NAN : (double)wsum[0] * (1.0 / 4611686018427387904.0);
    out[idx] = (float)((double)acc[idx] * back);
}

// The other regularisers of the coupling step (mcmc.cl:904-1243), one wave
// per q x q pair block; all read the block Jt = J + dJ and the model
// bimarginals f of the block and subtract a step from dJ:
//   SCADJ   SCAD penalty on the zero-gauge J0 (reg_SCADJ 904-932)
//   X       lambda_X |X| on X = sum Jt*C, C = f - fi fj (reg_X 1004-1028)
//   Xij     the same with one lambda per pair, lambdas[pair] (reg_Xij 972-1002)
//   SCADX   SCAD penalty on X (reg_SCADX 1077-1124)
//   expX    exp(-|X|/lam) penalty (reg_expX 1126-1150)
//   ddE     sign of every double difference J(a,b)-J(a',b)-J(a,b')+J(a',b')
//           (reg_ddE 1152-1187)
//   SCADddE SCAD-weighted double differences (reg_SCADddE 1191-1243)
// with fi, fj the row and column sums of f and
//   Xijab(c,d) = sum_ab Jt(a,b) (fi(a) - [a=c]) (fj(b) - [b=d])
//              = T - U(c) - V(d) + Jt(c,d),
//   U(c) = sum_b Jt(c,b) fj(b), V(d) = sum_a Jt(a,d) fi(a), T = sum_c fi(c) U(c)
// (the q^4 double loop of getXijab, mcmc.cl:1052-1067, in q^2 work).
enum RegMode { REG_SCADJ = 0, REG_X = 1, REG_XIJ = 2, REG_SCADX = 3, REG_EXPX = 4, REG_DDE = 5,
               REG_SCADDDE = 6 };

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(256) void k_reg_block(const float* __restrict__ bi,
                                                   const float* __restrict__ J,
                                                   float* __restrict__ dJ,
                                                   const float* __restrict__ lambdas, uint32_t nP,
                                                   uint32_t q, float gamma, float pc, float p0,
                                                   float p1, float p2, int mode) {
    __shared__ float sJ[4][32 * 32];
    __shared__ float sa[4][2][32];   // row / column quantities
    const uint32_t wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * 4 + wave;
    const uint32_t qq = q * q;
    const bool active = p < nP;   // wave-uniform
    float jt[16], fv[16], dj[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t e = l + 64 * r;
        jt[r] = fv[r] = dj[r] = 0.f;
        if (active && e < qq) {
            const size_t o = (size_t)p * qq + e;
            dj[r] = dJ[o];
            fv[r] = bi[o];
            jt[r] = J[o] + dj[r];
            sJ[wave][e] = jt[r];
        }
    }
    __syncthreads();
    if (mode == REG_DDE || mode == REG_SCADDDE) {
        const float lam = p0, rr = p1;
        if (active) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const uint32_t e = l + 64 * r;
                if (e < qq) {
                    const uint32_t a = e / q, b = e - a * q;
                    float dR = 0.f;
                    for (uint32_t g = 1; g < q; g++) {
                        const uint32_t a2 = (a + g) % q;
                        const float jr = sJ[wave][e] - sJ[wave][a2 * q + b];
                        for (uint32_t d = 1; d < q; d++) {
                            const uint32_t b2 = (b + d) % q;
                            const float ddE = jr + (sJ[wave][a2 * q + b2] - sJ[wave][a * q + b2]);
                            const float sg = fsign(ddE);
                            if (mode == REG_DDE) {
                                dR += sg;
                            } else {
                                const float ad = fabsf(ddE);
                                const float R = ad < rr ? 1.f
                                                : (ad < 4.f * rr ? (4.f * rr - ad) / (4.f * rr - rr)
                                                                 : 0.f);
                                dR += lam * sg * R;
                            }
                        }
                    }
                    dR = dR / (float)((q - 1) * (q - 1));
                    const float step = (mode == REG_DDE) ? lam * dR * gamma / (fv[r] + pc)
                                                         : gamma * dR / (fv[r] + pc);
                    dJ[(size_t)p * qq + e] = dj[r] - step;
                }
            }
        }
        return;
    }
    if (mode == REG_SCADJ) {
        // zero-gauge J0 of Jt: row means, column means, total mean
        if (active) {
            if (l < q) {
                float s = 0.f;
                for (uint32_t b = 0; b < q; b++) s += sJ[wave][l * q + b];
                sa[wave][0][l] = s / q;
            } else if (l >= 32 && l - 32 < q) {
                const uint32_t b = l - 32;
                float s = 0.f;
                for (uint32_t a = 0; a < q; a++) s += sJ[wave][a * q + b];
                sa[wave][1][b] = s / q;
            }
        }
        __syncthreads();
        if (!active) return;
        float mean = 0.f;
        for (uint32_t a = 0; a < q; a++) mean += sa[wave][0][a];
        mean /= q;
        const float lJ = p0, aa = p1;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t e = l + 64 * r;
            if (e < qq) {
                const uint32_t a = e / q, b = e - a * q;
                const float hi = sa[wave][0][a] - mean, hj = sa[wave][1][b] - mean;
                const float J0 = jt[r] - (hi + hj) - mean;
                const float aJ = fabsf(J0);
                float R = 0.f;
                if (aJ < lJ)
                    R = fminf(lJ, aJ / gamma);
                else if (aJ < aa * lJ)
                    R = (aa * lJ - aJ) / (aa - 1.f);
                R *= gamma * fsign(J0) / (fv[r] + pc);
                dJ[(size_t)p * qq + e] = dj[r] - R;
            }
        }
        return;
    }
    // X family: unimarginals of the block's bimarginals
    if (active) {
        if (l < q) {
            float s = 0.f;
            for (uint32_t b = 0; b < q; b++) s += bi[(size_t)p * qq + l * q + b];
            sa[wave][0][l] = s;
        } else if (l >= 32 && l - 32 < q) {
            const uint32_t b = l - 32;
            float s = 0.f;
            for (uint32_t a = 0; a < q; a++) s += bi[(size_t)p * qq + a * q + b];
            sa[wave][1][b] = s;
        }
    }
    __syncthreads();
    float C[16];
    float xs = 0.f, xn = 0.f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t e = l + 64 * r;
        C[r] = 0.f;
        if (active && e < qq) {
            const uint32_t a = e / q, b = e - a * q;
            C[r] = fv[r] - sa[wave][0][a] * sa[wave][1][b];
            xs += jt[r] * C[r];
            xn += C[r] * C[r] / (fv[r] + pc);
        }
    }
    const float X = wave_sum(xs), Xnorm = wave_sum(xn);
    const bool needXab = (mode == REG_SCADX || mode == REG_EXPX);
    __shared__ float sUV[4][2][32];
    float T = 0.f;
    if (needXab) {
        if (active) {
            if (l < q) {
                float u = 0.f;
                for (uint32_t b = 0; b < q; b++) u += sJ[wave][l * q + b] * sa[wave][1][b];
                sUV[wave][0][l] = u;
            } else if (l >= 32 && l - 32 < q) {
                const uint32_t d = l - 32;
                float v = 0.f;
                for (uint32_t a = 0; a < q; a++) v += sJ[wave][a * q + d] * sa[wave][0][a];
                sUV[wave][1][d] = v;
            }
        }
        __syncthreads();
        if (active)
            for (uint32_t c = 0; c < q; c++) T += sa[wave][0][c] * sUV[wave][0][c];
    }
    if (!active) return;
    const float sX = fsign(X), aX = fabsf(X);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t e = l + 64 * r;
        if (e >= qq) continue;
        const uint32_t a = e / q, b = e - a * q;
        const float f = fv[r];
        float out;
        if (mode == REG_X || mode == REG_XIJ) {
            const float lX0 = (mode == REG_X) ? p0 : lambdas[p];
            const float lX = fminf(lX0, aX / (Xnorm * gamma));
            out = (mode == REG_X) ? dj[r] - gamma * lX * C[r] * sX / (f + pc)
                                  : dj[r] - gamma * (lX * C[r] * sX) / (f + pc);
        } else {
            const float Xab = T - sUV[wave][0][a] - sUV[wave][1][b] + jt[r];
            const float grad = C[r] + f * (X - Xab);
            if (mode == REG_SCADX) {
                const float s = p0, rr0 = p1, aa = p2;
                float R = 0.f;
                if (aX < rr0) {
                    const float rr = fminf(rr0, aX / (Xnorm * gamma * s));
                    R = rr * sX;
                } else if (aX < aa * rr0) {
                    R = (aa * rr0 * sX - X) / (aa - 1.f);
                }
                R *= gamma * s * grad / (f + pc);
                out = dj[r] - R;
            } else {  // expX
                const float R = sX * expf(-aX / p0) * grad;
                out = dj[r] - gamma * R / (f + pc);
            }
        }
        dJ[(size_t)p * qq + e] = out;
    }
}

}  // namespace mi3
