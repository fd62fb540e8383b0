// Learned proposal: a masked autoregressive MLP over sequence positions
// (MADE, Germain et al. 2015) used as an independence Metropolis-Hastings
// proposal beside the single-site sweep.
//
// Network (L sites, q letters, H hidden units per layer):
//   h1 = relu(b1 + sum_i W1[:, q*i + s_i])          (one-hot input: a gather-sum)
//   h2 = relu(W2 h1 + b2)
//   logits = W3 h2 + b3,  p(s_j | s_<j) = softmax over the q logits of site j
// with MADE masks baked into the weights (site j's logits see only s_<j).
// GEMMs run in bf16 on the matrix cores with float32 accumulation; the
// one-hot layer, softmax and sampling run in float32.  Sequences are
// site-major (L rows of `ld` bytes), the engine's walker layout.
#include "common.h"

namespace mi3 {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// C = alpha * op(A) op(B) + beta * C + bias[col], optionally relu; fp32 C
// and/or bf16 Cb.  op(A) is M x K (A row-major M x K, or K x M if TA),
// op(B) is K x N (B row-major K x N, or N x K if TB).  Work-group tile
// 128 x 128 (4 waves of 64 x 64 = 2 x 2 v_mfma_f32_32x32x16_bf16), K step
// 32 staged through LDS as [row][k] for A and [col][k] for B.  Requires
// M, N, K, lda, ldb multiples of 8 and 16-byte aligned A, B (host-checked).
constexpr int GT = 128, GK = 32, GKP = GK + 8;   // tile, k step, padded LDS row

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, float alpha,
                                                   const bf16* __restrict__ A, int lda,
                                                   const bf16* __restrict__ B, int ldb,
                                                   float beta, float* __restrict__ C, int ldc,
                                                   bf16* __restrict__ Cb, int ldcb,
                                                   const float* __restrict__ bias, int relu) {
    __shared__ __attribute__((aligned(16))) bf16 As[GT][GKP];
    __shared__ __attribute__((aligned(16))) bf16 Bs[GT][GKP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 31, lg = lane >> 5;
    f32x16 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[x][y][e] = 0.f;
    const uint4 zero = make_uint4(0, 0, 0, 0);

    for (int k0 = 0; k0 < K; k0 += GK) {
        // stage A' (GT rows x GK k) and B' (GK k x GT cols): 512 chunks of
        // 8 bf16 each, two per thread
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int c = tid + 256 * r;
            if (!TA) {
                const int m = c >> 2, kk = (c & 3) * 8;
                uint4 v = zero;
                if (m0 + m < M && k0 + kk < K)
                    v = *reinterpret_cast<const uint4*>(A + (size_t)(m0 + m) * lda + k0 + kk);
                *reinterpret_cast<uint4*>(&As[m][kk]) = v;
            } else {
                const int k = c >> 4, mm = (c & 15) * 8;
                uint4 v = zero;
                if (k0 + k < K && m0 + mm < M)
                    v = *reinterpret_cast<const uint4*>(A + (size_t)(k0 + k) * lda + m0 + mm);
                const bf16* e8 = reinterpret_cast<const bf16*>(&v);
#pragma unroll
                for (int e = 0; e < 8; e++) As[mm + e][k] = e8[e];
            }
            if (TB) {
                const int n = c >> 2, kk = (c & 3) * 8;
                uint4 v = zero;
                if (n0 + n < N && k0 + kk < K)
                    v = *reinterpret_cast<const uint4*>(B + (size_t)(n0 + n) * ldb + k0 + kk);
                *reinterpret_cast<uint4*>(&Bs[n][kk]) = v;
            } else {
                const int k = c >> 4, nn = (c & 15) * 8;
                uint4 v = zero;
                if (k0 + k < K && n0 + nn < N)
                    v = *reinterpret_cast<const uint4*>(B + (size_t)(k0 + k) * ldb + n0 + nn);
                const bf16* e8 = reinterpret_cast<const bf16*>(&v);
#pragma unroll
                for (int e = 0; e < 8; e++) Bs[nn + e][k] = e8[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int x = 0; x < 2; x++)
                a[x] = *reinterpret_cast<const bf16x8*>(&As[64 * wm + 32 * x + lr][kk + 8 * lg]);
#pragma unroll
            for (int y = 0; y < 2; y++)
                b[y] = *reinterpret_cast<const bf16x8*>(&Bs[64 * wn + 32 * y + lr][kk + 8 * lg]);
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x], b[y], acc[x][y], 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: value e of lane l is row 8*(e/4) + 4*(l/32) + e%4,
    // column l%32
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            const int col = n0 + 64 * wn + 32 * y + lr;
            if (col >= N) continue;
            const float bc = bias ? bias[col] : 0.f;
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = m0 + 64 * wm + 32 * x + 8 * (e >> 2) + 4 * lg + (e & 3);
                if (row >= M) continue;
                float v = alpha * acc[x][y][e] + bc;
                if (beta != 0.f) v += beta * C[(size_t)row * ldc + col];
                if (relu) v = fmaxf(v, 0.f);
                if (C) C[(size_t)row * ldc + col] = v;
                if (Cb) Cb[(size_t)row * ldcb + col] = (bf16)v;
            }
        }
}

// h1[b][k] = b1[k] + sum_i W1t[(q*i + s_i)*H + k]  (W1t: (L*q, H), masked);
// h1b = bf16(relu(h1)).  One thread per (b, k), k fastest.
__global__ void k_made_embed(const uint8_t* __restrict__ seq, uint32_t ld, uint32_t B, uint32_t L,
                             uint32_t q, uint32_t H, const float* __restrict__ W1t,
                             const float* __restrict__ b1, float* __restrict__ h1,
                             bf16* __restrict__ h1b) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)B * H) return;
    const uint32_t b = (uint32_t)(idx / H), k = (uint32_t)(idx - (uint64_t)b * H);
    float v = b1[k];
    for (uint32_t i = 0; i < L; i++) v += W1t[(size_t)(q * i + seq[(size_t)i * ld + b]) * H + k];
    h1[idx] = v;
    h1b[idx] = (bf16)fmaxf(v, 0.f);
}

// dW1t[(q*i + a)*H + k] = sum over b with s_i(b) == a of dh1[b][k]: one
// work-group per (site i, 256 units); thread k keeps its q sums in an LDS
// column of its own and walks the batch in order, adding dh1[b][k] into the
// row of the letter s_i(b) (a wave-uniform scalar load), so each sequence
// costs one LDS read and write instead of a select over every letter.
// Writes every entry, so no zeroing is needed; deterministic.
constexpr int EG_T = 256, EG_QMAX = 25;
__global__ __launch_bounds__(EG_T) void k_made_embed_grad(const uint8_t* __restrict__ seq,
                                                          uint32_t ld, uint32_t B, uint32_t L,
                                                          uint32_t q, uint32_t H,
                                                          const float* __restrict__ dh1,
                                                          float* __restrict__ dW1t) {
    __shared__ float acc[EG_QMAX][EG_T];
    const uint32_t nkb = (H + EG_T - 1) / EG_T;
    const uint32_t i = blockIdx.x / nkb, k = (blockIdx.x % nkb) * EG_T + threadIdx.x;
    for (uint32_t a = 0; a < q; a++) acc[a][threadIdx.x] = 0.f;
    const uint8_t* si = seq + (size_t)i * ld;
    if (k < H) {
        // eight sequences per step: their letters and gradients are loaded
        // before the eight additions (in batch order)
        constexpr int U = 8;
        uint32_t b = 0;
        for (; b + U <= B; b += U) {
            uint32_t sv[U];
            float gv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                sv[u] = si[b + u];
                gv[u] = dh1[(size_t)(b + u) * H + k];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t s = __builtin_amdgcn_readfirstlane(sv[u]);
                acc[s < q ? s : q - 1][threadIdx.x] += gv[u];
            }
        }
        for (; b < B; b++) {
            const uint32_t s = __builtin_amdgcn_readfirstlane(si[b]);
            acc[s < q ? s : q - 1][threadIdx.x] += dh1[(size_t)b * H + k];
        }
        for (uint32_t a = 0; a < q; a++) dW1t[(size_t)(q * i + a) * H + k] = acc[a][threadIdx.x];
    }
}

// Per sequence b (one work-group): nll[b] = -sum_j log softmax(logits_j)[s_j]
// and dlogits = gscale * (softmax - onehot), written bf16 (and fp32 if
// dlf != null).  Threads stride over the sites; the block sum is a fixed
// tree, so nll is deterministic.
__global__ __launch_bounds__(256) void k_made_xent(const float* __restrict__ logits,
                                                   const uint8_t* __restrict__ seq, uint32_t ld,
                                                   uint32_t B, uint32_t L, uint32_t q, float gscale,
                                                   float* __restrict__ nll, bf16* __restrict__ dlb,
                                                   float* __restrict__ dlf) {
    __shared__ float red[256];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const size_t row = (size_t)b * L * q;
    float part = 0.f;
    for (uint32_t j = tid; j < L; j += 256) {
        const float* z = logits + row + (size_t)j * q;
        float mx = -INFINITY;
        for (uint32_t a = 0; a < q; a++) mx = fmaxf(mx, z[a]);
        float se = 0.f;
        for (uint32_t a = 0; a < q; a++) se += expf(z[a] - mx);
        const float lse = mx + logf(se);
        const uint32_t s = seq[(size_t)j * ld + b];
        part += lse - z[s];
        if (dlb || dlf) {
            for (uint32_t a = 0; a < q; a++) {
                const float g = gscale * (expf(z[a] - lse) - (a == s ? 1.f : 0.f));
                if (dlb) dlb[row + (size_t)j * q + a] = (bf16)g;
                if (dlf) dlf[row + (size_t)j * q + a] = g;
            }
        }
    }
    red[tid] = part;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    if (tid == 0) nll[b] = red[0];
}

// out = dh * (hpre > 0) (fp32) and its bf16 copy.
__global__ void k_relu_grad(uint64_t n, const float* __restrict__ dh,
                            const float* __restrict__ hpre, float* __restrict__ out,
                            bf16* __restrict__ outb) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = hpre[i] > 0.f ? dh[i] : 0.f;
    if (out) out[i] = v;
    if (outb) outb[i] = (bf16)v;
}

// Fused autoregressive pass over all L sites for tiles of MF_T = 16 walkers,
// one work-group of 4 waves per tile (several tiles share a CU): with
// draw != 0 it samples s_j by inversion of u = Philox((j, walker_base + b,
// counter, TAG_MADE)) and writes it to row j of seq; with draw == 0 it reads
// s_j from seq (teacher forcing).  Either way nll[b] = -sum_j log p(s_j).
//
// The degree masks make each hidden unit final once: unit k (degree
// m_k = 1 + k mod (L-1)) of layer 1 sees sites i < m_k, unit k of layer 2
// sees layer-1 units of degree <= m_k, and site j's logits see layer-2 units
// of degree <= j.  So at site j the layer-1 units of degree j have all their
// inputs, the layer-2 units of degree j can be formed from them, and the
// logits need nothing newer: each unit is computed once per walker, at its
// degree, instead of the whole hidden layer at every site (L times less
// matrix-core work than a GEMM per site).
//
// Per site j: (A) layer-1 units of degree j, kept as float32 sums in
// registers (thread t owns units t and t + 256 for the tile's walkers), go
// to LDS as bf16(relu); (B) layer-2 units of degree j: [16 walkers x H] x W2b
// rows of those units on v_mfma_f32_16x16x32_bf16, the k-steps dealt to the
// 4 waves (wave 0 starts from the bias) and the partials summed in a fixed
// order, then bf16(relu(.)) into LDS; (C) site j's q logits the same way
// (two column tiles) from the layer-2 units, W3b rows q*j..q*j+q-1 and b3;
// (D) 16 lanes per walker: softmax, draw or read, nll, then wave 1 draws the
// next site's uniforms; (E) the float32 layer-1 sums of units of degree
// > j add W1t's row of the site's letter (the sums are formed in site order,
// as the teacher-forced embedding forms them).
//
// The B operands and biases of (B) and (C) for site j+1 are loaded into
// registers right after the products of site j use them; when teacher
// forcing, so are the next site's letters.  Requires H % 32 == 0,
// H <= 512, q <= 25.
constexpr int MF_T = 16, MF_WAVES = 4, MF_THREADS = 64 * MF_WAVES, MF_HMAX = 512;
constexpr int MF_HU = MF_HMAX / MF_THREADS;     // layer-1 units per thread
constexpr int MF_QMAX = 25;
constexpr int MF_HP = MF_HMAX + 8;              // padded LDS row (bf16) of a walker's units
constexpr int MF_KS = MF_HMAX / 32 / MF_WAVES;  // k-steps of 32 per wave
constexpr int MF_NC = 32;                       // logit columns (two tiles of 16)
static_assert(MF_THREADS == 16 * MF_T, "16 lanes per walker in the softmax");

struct MadeFusedLds {
    bf16 h1[MF_T][MF_HP];                  // bf16(relu(layer 1)), final units only
    bf16 h2[MF_T][MF_HP];                  // bf16(relu(layer 2)), final units only
    float part[MF_WAVES][MF_T][MF_NC + 1]; // per-wave partial products
    float z[MF_T][MF_NC + 1];              // the site's logits
    float u[2][MF_T];                      // uniforms of this / the next site
    uint32_t s[MF_T];                      // the site's letters
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

// B operand of one wave's k-steps for one 16-column tile: this lane's 8
// k-values of column lane%16 (zero for a negative row or k past H), and for
// wave 0 the column's bias (the accumulator's start).
struct MadeB {
    bf16x8 b[MF_KS];
    float bias;
};

__device__ __forceinline__ void made_load_b(MadeB& r, const bf16* __restrict__ W,
                                            const float* __restrict__ bias, uint32_t H, int row,
                                            int lg, int wave) {
    const bf16x8 zero8 = {};
#pragma unroll
    for (int t = 0; t < MF_KS; t++) {
        const uint32_t k0 = 32 * (wave + MF_WAVES * t);
        r.b[t] = (row >= 0 && k0 < H)
            ? *reinterpret_cast<const bf16x8*>(W + (size_t)row * H + k0 + 8 * lg) : zero8;
    }
    r.bias = (wave == 0 && row >= 0) ? bias[row] : 0.f;
}

// part[wave][.][c0 + .] = bias (wave 0) + X[16 x H] . (this wave's k-steps of b)
__device__ __forceinline__ void made_product(const bf16 (*X)[MF_HP], const MadeB& b, uint32_t H,
                                             float (*part)[MF_NC + 1], int c0, int lr, int lg,
                                             int wave) {
    f32x4 acc;
#pragma unroll
    for (int e = 0; e < 4; e++) acc[e] = b.bias;
#pragma unroll
    for (int t = 0; t < MF_KS; t++) {
        const uint32_t k0 = 32 * (wave + MF_WAVES * t);
        if (k0 < H) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&X[lr][k0 + 8 * lg]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b.b[t], acc, 0, 0, 0);
        }
    }
    // D layout: value e of lane l is row 4*(l/16) + e, column l%16
#pragma unroll
    for (int e = 0; e < 4; e++) part[4 * lg + e][c0 + lr] = acc[e];
}

__device__ __forceinline__ float made_part_sum(const float (*part)[MF_T][MF_NC + 1], uint32_t r,
                                               uint32_t c) {
    return ((part[0][r][c] + part[1][r][c]) + part[2][r][c]) + part[3][r][c];
}

__global__ __launch_bounds__(MF_THREADS) void k_made_fused(
    uint32_t L, uint32_t q, uint32_t H, uint32_t B, const float* __restrict__ W1t,
    const float* __restrict__ b1, const bf16* __restrict__ W2b, const float* __restrict__ b2,
    const bf16* __restrict__ W3b, const float* __restrict__ b3, uint32_t k0, uint32_t k1,
    uint32_t counter, uint32_t walker_base, uint8_t* __restrict__ seq, uint32_t ld,
    float* __restrict__ nll, int draw) {
    __shared__ MadeFusedLds sm;
    const int tid = threadIdx.x, lane = tid & 63;
    const int lr = lane & 15, lg = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t tile0 = blockIdx.x * MF_T;
    const uint32_t Lm1 = L - 1;
    // units of degree jj: k = (jj - 1) + c (L - 1), c = 0, 1, ...
    auto n_units = [&](uint32_t jj) -> uint32_t {
        return (jj >= 1 && jj < L && jj - 1 < H) ? (H - (jj - 1) + Lm1 - 1) / Lm1 : 0u;
    };
    uint32_t deg[MF_HU];
    float h1r[MF_HU][MF_T];
#pragma unroll
    for (int u = 0; u < MF_HU; u++) {
        const uint32_t k = tid + MF_THREADS * u;
        deg[u] = k < H ? 1 + k % Lm1 : 0xffffffffu;
        const float v = k < H ? b1[k] : 0.f;
#pragma unroll
        for (int w = 0; w < MF_T; w++) h1r[u][w] = v;
    }
    for (uint32_t e = tid; e < MF_T * MF_HP; e += MF_THREADS) {
        (&sm.h1[0][0])[e] = (bf16)0.f;
        (&sm.h2[0][0])[e] = (bf16)0.f;
    }
    MadeB w2r, w3r[2];
    auto load_w2 = [&](uint32_t jj) {
        const uint32_t nU = n_units(jj);
        const int row = (nU <= MF_T && (uint32_t)lr < nU) ? (int)((jj - 1) + lr * Lm1) : -1;
        made_load_b(w2r, W2b, b2, H, row, lg, wave);
    };
    auto load_w3 = [&](uint32_t jj) {
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const uint32_t c = 16 * t + lr;
            made_load_b(w3r[t], W3b, b3, H, (jj < L && c < q) ? (int)(q * jj + c) : -1, lg, wave);
        }
    };
    // uniforms of site jj for the tile's walkers (wave 1, lanes < 16)
    auto draw_u = [&](uint32_t jj) {
        if (draw && wave == 1 && lane < MF_T && jj < L) {
            const uint4 r = philox4x32_10(make_uint4(jj, walker_base + tile0 + lane, counter,
                                                     TAG_MADE), k0, k1);
            sm.u[jj & 1][lane] = (float)(r.x >> 8) * 0x1.0p-24f;
        }
    };
    // teacher forcing: every lane of walker w's group holds its letter at the
    // next site
    uint32_t s_next = 0;
    auto load_s = [&](uint32_t jj) {
        const uint32_t b = tile0 + ((uint32_t)tid >> 4);
        s_next = (!draw && b < B && jj < L) ? seq[(size_t)jj * ld + b] : 0u;
    };
    load_w2(0);
    load_w3(0);
    load_s(0);
    draw_u(0);
    float nacc = 0.f;    // nll of walker tile0 + tid / 16 (every lane of its group)

    for (uint32_t j = 0; j < L; j++) {
        // (A) layer-1 units of degree j are final
#pragma unroll
        for (int u = 0; u < MF_HU; u++) {
            if (deg[u] == j) {
                const uint32_t k = tid + MF_THREADS * u;
#pragma unroll
                for (int w = 0; w < MF_T; w++) sm.h1[w][k] = (bf16)fmaxf(h1r[u][w], 0.f);
            }
        }
        __syncthreads();
        // (B) layer-2 units of degree j, 16 at a time
        const uint32_t nU = n_units(j);
        for (uint32_t c0 = 0; c0 < nU; c0 += MF_T) {
            if (nU > MF_T) {   // more than one column tile: operands loaded here
                const uint32_t c = c0 + lr;
                made_load_b(w2r, W2b, b2, H, c < nU ? (int)((j - 1) + c * Lm1) : -1, lg, wave);
            }
            made_product(sm.h1, w2r, H, sm.part[wave], 0, lr, lg, wave);
            __syncthreads();
            {
                const uint32_t r = tid / MF_T, c = tid % MF_T;   // 256 = 16 x 16 entries
                if (c0 + c < nU)
                    sm.h2[r][(j - 1) + (c0 + c) * Lm1] =
                        (bf16)fmaxf(made_part_sum(sm.part, r, c), 0.f);
            }
            __syncthreads();
        }
        load_w2(j + 1);
        // (C) logits of site j (columns a < q), b3 included
        made_product(sm.h2, w3r[0], H, sm.part[wave], 0, lr, lg, wave);
        made_product(sm.h2, w3r[1], H, sm.part[wave], 16, lr, lg, wave);
        load_w3(j + 1);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t e = tid + MF_THREADS * h, r = e / MF_NC, c = e % MF_NC;
            if (c < q) sm.z[r][c] = made_part_sum(sm.part, r, c);
        }
        __syncthreads();
        // (D) softmax and the letter: 16 lanes per walker (walker tid / 16;
        // lane l holds letters l and l + 16), max, sum and the inverse-CDF
        // search by 16-lane shuffles and one ballot; wave 1 then draws the
        // next site's uniforms
        {
            const uint32_t w = (uint32_t)tid >> 4, l = (uint32_t)tid & 15;
            const float z1 = (l < q) ? sm.z[w][l] : -INFINITY;
            const float z2 = (l + 16 < q) ? sm.z[w][l + 16] : -INFINITY;
            float mx = fmaxf(z1, z2);
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 16));
            float c1 = (l < q) ? __expf(z1 - mx) : 0.f;
            float c2 = (l + 16 < q) ? __expf(z2 - mx) : 0.f;
            // inclusive prefix sums over the group (letters in order)
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const float t1 = __shfl_up(c1, o, 16), t2 = __shfl_up(c2, o, 16);
                if (l >= (uint32_t)o) {
                    c1 += t1;
                    c2 += t2;
                }
            }
            const float tot1 = __shfl(c1, 15, 16);
            const float se = tot1 + __shfl(c2, 15, 16);
            uint32_t s;
            if (draw) {
                // first letter a (a + 1 < q) whose cumulative mass exceeds u
                const float uu = sm.u[j & 1][w] * se;
                const uint64_t m1 = __ballot((l + 1 < q) && uu < c1);
                const uint64_t m2 = __ballot((l + 17 < q) && uu < tot1 + c2);
                const uint32_t sh = (uint32_t)tid & 48u;
                const uint32_t g1 = (uint32_t)(m1 >> sh) & 0xffffu;
                const uint32_t g2 = (uint32_t)(m2 >> sh) & 0xffffu;
                s = g1 ? (uint32_t)__builtin_ctz(g1)
                       : (g2 ? 16u + (uint32_t)__builtin_ctz(g2) : q - 1);
                if (l == 0 && tile0 + w < B) seq[(size_t)j * ld + tile0 + w] = (uint8_t)s;
            } else {
                s = s_next < q ? s_next : q - 1;
            }
            nacc += mx + __logf(se) - sm.z[w][s];
            if (l == 0) sm.s[w] = s;
        }
        draw_u(j + 1);
        load_s(j + 1);
        __syncthreads();
        // (E) sums of the layer-1 units that see site j
        if (j + 1 < L) {
            uint32_t sw[MF_T];
#pragma unroll
            for (int w = 0; w < MF_T; w++) sw[w] = sm.s[w];
#pragma unroll
            for (int u = 0; u < MF_HU; u++) {
                if (deg[u] != 0xffffffffu && deg[u] > j) {
                    const uint32_t k = tid + MF_THREADS * u;
                    float v[MF_T];
#pragma unroll
                    for (int w = 0; w < MF_T; w++) v[w] = W1t[(size_t)(q * j + sw[w]) * H + k];
#pragma unroll
                    for (int w = 0; w < MF_T; w++) h1r[u][w] += v[w];
                }
            }
        }
    }
    if ((tid & 15) == 0 && tile0 + (tid >> 4) < B) nll[tile0 + (tid >> 4)] = nacc;
}

// Independence Metropolis-Hastings: walker w moves to its proposal with
// probability min(1, exp(-(Ep - Ec) + log q(cur) - log q(prop))), with the
// proposal's negative log-likelihoods nc = -log q(cur), np = -log q(prop);
// u from Philox((counter lo, hi, walker, TAG_MHAC)).  Accepted columns are
// copied (site-major, L rows), with the energy and nll.
__global__ void k_mh_accept(uint32_t N, uint32_t L, uint8_t* __restrict__ cur, uint32_t ldc,
                            const uint8_t* __restrict__ prop, uint32_t ldp,
                            float* __restrict__ Ec, const float* __restrict__ Ep,
                            float* __restrict__ nc, const float* __restrict__ np, uint32_t k0,
                            uint32_t k1, uint64_t counter, uint32_t walker_base,
                            uint32_t* __restrict__ naccept) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= N) return;
    const uint4 r = philox4x32_10(
        make_uint4((uint32_t)counter, (uint32_t)(counter >> 32), walker_base + w, TAG_MHAC), k0,
        k1);
    const float u = ((float)(r.x >> 8) + 0.5f) * 0x1.0p-24f;
    const float logr = -(Ep[w] - Ec[w]) - nc[w] + np[w];
    if (logf(u) < logr) {
        for (uint32_t i = 0; i < L; i++) cur[(size_t)i * ldc + w] = prop[(size_t)i * ldp + w];
        Ec[w] = Ep[w];
        nc[w] = np[w];
        atomicAdd(naccept, 1u);
    }
}

// Adam step on float32 parameters p with gradient g (times the 0/1 mask if
// given): m, v moments; bias-corrected with step t >= 1.  Optionally writes
// the bf16 copy pb used by the matrix-core products.
__global__ void k_adam(uint64_t n, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ mask, float lr, float b1, float b2, float eps,
                       float c1, float c2, bf16* __restrict__ pb) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    if (mask) gi *= mask[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    if (mask) pi *= mask[i];
    p[i] = pi;
    if (pb) pb[i] = (bf16)pi;
}

// out[c] = sum over rows r of x[r][c] (rows x cols row-major).  A work-group
// takes 32 columns with 8 row lanes each: row lane t sums rows t, t + 8, ...
// in order (each load instruction reads 8 rows x 128 contiguous bytes), then
// the 8 partial sums are added in a fixed order, so the result is
// deterministic.
constexpr int CS_COLS = 32, CS_ROWL = 8;
__global__ __launch_bounds__(CS_COLS * CS_ROWL) void k_colsum(uint32_t rows, uint32_t cols,
                                                              const float* __restrict__ x,
                                                              float* __restrict__ out) {
    __shared__ float part[CS_ROWL][CS_COLS + 1];
    const uint32_t cl = threadIdx.x % CS_COLS, t = threadIdx.x / CS_COLS;
    const uint32_t c = blockIdx.x * CS_COLS + cl;
    float s = 0.f;
    if (c < cols)
        for (uint32_t r = t; r < rows; r += CS_ROWL) s += x[(size_t)r * cols + c];
    part[t][cl] = s;
    __syncthreads();
    if (t == 0 && c < cols) {
        float v = part[0][cl];
#pragma unroll
        for (int k = 1; k < CS_ROWL; k++) v += part[k][cl];
        out[c] = v;
    }
}

// pb = bf16(p).
__global__ void k_to_bf16(uint64_t n, const float* __restrict__ p, bf16* __restrict__ pb) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = (bf16)p[i];
}

}  // namespace mi3
