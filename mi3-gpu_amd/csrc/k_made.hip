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
// thread per (site i, unit k) holds the q sums in registers and walks the
// batch (the site's residues are read as a broadcast, dh1 coalesced along
// k).  Writes every entry, so no zeroing is needed; deterministic.
__global__ __launch_bounds__(256) void k_made_embed_grad(const uint8_t* __restrict__ seq,
                                                         uint32_t ld, uint32_t B, uint32_t L,
                                                         uint32_t q, uint32_t H,
                                                         const float* __restrict__ dh1,
                                                         float* __restrict__ dW1t) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)L * H) return;
    const uint32_t i = (uint32_t)(idx / H), k = (uint32_t)(idx - (uint64_t)i * H);
    float acc[32];
#pragma unroll
    for (uint32_t a = 0; a < 32; a++) acc[a] = 0.f;
    const uint8_t* si = seq + (size_t)i * ld;
    for (uint32_t b = 0; b < B; b++) {
        const uint32_t s = si[b];
        const float g = dh1[(size_t)b * H + k];
#pragma unroll
        for (uint32_t a = 0; a < 32; a++) acc[a] += (a == s) ? g : 0.f;
    }
    for (uint32_t a = 0; a < q; a++) {
        float v = 0.f;
#pragma unroll
        for (uint32_t e = 0; e < 32; e++)
            if (e == a) v = acc[e];
        dW1t[(size_t)(q * i + a) * H + k] = v;
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

// Autoregressive sampling, site j: logits_j = W3[q*j .. q*j+q) h2 + b3,
// s_j ~ softmax(logits_j) by inversion of u = Philox((j, walker, counter,
// TAG_MADE)) -- one thread per walker b;
// nll[b] -= log p(s_j) (written at j = 0, accumulated after); s_j written to
// row j of the site-major output.
// The site's q rows of W3 (bf16) are staged in LDS.
__global__ __launch_bounds__(256) void k_made_site(uint32_t j, const bf16* __restrict__ h2b,
                                                   uint32_t H, const bf16* __restrict__ W3b,
                                                   const float* __restrict__ b3, uint32_t q,
                                                   uint32_t B, uint32_t k0, uint32_t k1,
                                                   uint32_t counter, uint32_t walker_base,
                                                   uint8_t* __restrict__ sout, uint32_t ld,
                                                   float* __restrict__ nll) {
    extern __shared__ float w3s[];   // q x H floats
    for (uint32_t e = threadIdx.x; e < q * H; e += blockDim.x)
        w3s[e] = (float)W3b[(size_t)q * j * H + e];
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float z[32];
#pragma unroll
    for (uint32_t a = 0; a < 32; a++) z[a] = 0.f;
    const bf16* h = h2b + (size_t)b * H;
    for (uint32_t k = 0; k < H; k += 8) {
        const bf16x8 hv = *reinterpret_cast<const bf16x8*>(h + k);
        float hf[8];
#pragma unroll
        for (int e = 0; e < 8; e++) hf[e] = (float)hv[e];
#pragma unroll
        for (uint32_t a = 0; a < 32; a++) {
            if (a < q) {
                float acc = 0.f;
#pragma unroll
                for (int e = 0; e < 8; e++) acc += hf[e] * w3s[a * H + k + e];
                z[a] += acc;
            }
        }
    }
    float mx = -INFINITY;
#pragma unroll
    for (uint32_t a = 0; a < 32; a++)
        if (a < q) {
            z[a] += b3[q * j + a];
            mx = fmaxf(mx, z[a]);
        }
    float se = 0.f;
#pragma unroll
    for (uint32_t a = 0; a < 32; a++)
        if (a < q) se += expf(z[a] - mx);
    const uint4 r = philox4x32_10(make_uint4(j, walker_base + b, counter, TAG_MADE), k0, k1);
    const float u = (float)(r.x >> 8) * 0x1.0p-24f * se;
    // first a whose cumulative mass exceeds u (the last residue otherwise)
    float c = 0.f;
    uint32_t s = q - 1;
    for (uint32_t a = 0; a + 1 < q; a++) {
        c += expf(z[a] - mx);
        if (u < c) {
            s = a;
            break;
        }
    }
    float zs = z[0];
#pragma unroll
    for (uint32_t a = 0; a < 32; a++)
        if (a == s) zs = z[a];
    const float t = mx + logf(se) - zs;
    nll[b] = j == 0 ? t : nll[b] + t;
    sout[(size_t)j * ld + b] = (uint8_t)s;
}

// After sampling site j: h1[b][:] += W1t[(q*j + s_j)*H + :], h1b = bf16(relu(h1)).
__global__ void k_made_h1_update(uint32_t j, const uint8_t* __restrict__ sout, uint32_t ld,
                                 uint32_t B, uint32_t q, uint32_t H,
                                 const float* __restrict__ W1t, float* __restrict__ h1,
                                 bf16* __restrict__ h1b) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)B * H) return;
    const uint32_t b = (uint32_t)(idx / H), k = (uint32_t)(idx - (uint64_t)b * H);
    const float v = h1[idx] + W1t[(size_t)(q * j + sout[(size_t)j * ld + b]) * H + k];
    h1[idx] = v;
    h1b[idx] = (bf16)fmaxf(v, 0.f);
}

// h1 = b1 (start of sampling), h1b = bf16(relu(b1)).
__global__ void k_made_h1_init(uint32_t B, uint32_t H, const float* __restrict__ b1,
                               float* __restrict__ h1, bf16* __restrict__ h1b) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)B * H) return;
    const float v = b1[idx % H];
    h1[idx] = v;
    h1b[idx] = (bf16)fmaxf(v, 0.f);
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

// out[c] = sum over rows r of x[r][c] (rows x cols row-major), one thread
// per column, rows in order (deterministic).
__global__ void k_colsum(uint32_t rows, uint32_t cols, const float* __restrict__ x,
                         float* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (uint32_t r = 0; r < rows; r++) s += x[(size_t)r * cols + c];
    out[c] = s;
}

// pb = bf16(p).
__global__ void k_to_bf16(uint64_t n, const float* __restrict__ p, bf16* __restrict__ pb) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = (bf16)p[i];
}

}  // namespace mi3
