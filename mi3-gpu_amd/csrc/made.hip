// libmi3hip.so: C ABI of the learned-proposal kernels (k_made.hip).  The
// launchers take raw device pointers and a HIP stream, so that the caller's
// own allocator and stream (the training step's framework) own the memory.
#include <hip/hip_runtime.h>

#include <string>

#include "k_made.hip"
#include "launch.h"
#include "mi3hip.h"

using namespace mi3;

namespace {
thread_local std::string g_made_error;

int mfail(int code, const std::string& m) {
    g_made_error = m;
    return code;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t blocks(uint64_t n, uint32_t bs = 256) { return (uint32_t)((n + bs - 1) / bs); }

int launched() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MI3_OK : mfail(MI3_EHIP, hipGetErrorString(e));
}

bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }
}  // namespace

extern "C" {

const char* mi3_made_last_error(void) { return g_made_error.c_str(); }

int mi3_gemm_bf16(void* stream, int transA, int transB, int M, int N, int K, float alpha,
                  const void* A, int lda, const void* B, int ldb, float beta, float* C, int ldc,
                  void* Cb, int ldcb, const float* bias, int relu) {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || (!C && !Cb))
        return mfail(MI3_EINVAL, "gemm: bad sizes or null operands");
    if (M % 8 || N % 8 || K % 8 || lda % 8 || ldb % 8 || !al16(A) || !al16(B))
        return mfail(MI3_EINVAL, "gemm: M, N, K, lda, ldb must be multiples of 8, A and B "
                                 "16-byte aligned");
    if ((!transA && lda < K) || (transA && lda < M) || (!transB && ldb < N) ||
        (transB && ldb < K) || (C && ldc < N) || (Cb && ldcb < N))
        return mfail(MI3_EINVAL, "gemm: leading dimension too small");
    if (beta != 0.f && !C) return mfail(MI3_EINVAL, "gemm: beta needs C");
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    const bf16* a = static_cast<const bf16*>(A);
    const bf16* b = static_cast<const bf16*>(B);
    bf16* cb = static_cast<bf16*>(Cb);
#define MI3_G(TA_, TB_)                                                                       \
    hipLaunchKernelGGL((k_gemm_bf16<TA_, TB_>), grid, dim3(256), 0, S(stream), M, N, K, alpha, \
                       a, lda, b, ldb, beta, C, ldc, cb, ldcb, bias, relu)
    if (!transA && !transB) MI3_G(false, false);
    else if (!transA && transB) MI3_G(false, true);
    else if (transA && !transB) MI3_G(true, false);
    else MI3_G(true, true);
#undef MI3_G
    return launched();
}

int mi3_made_embed(void* stream, const uint8_t* seq, uint32_t ld, uint32_t B, uint32_t L,
                   uint32_t q, uint32_t H, const float* W1t, const float* b1, float* h1,
                   void* h1b) {
    if (!seq || !W1t || !b1 || !h1 || !h1b || ld < B || q < 2 || q > MAX_Q)
        return mfail(MI3_EINVAL, "made_embed: bad arguments");
    hipLaunchKernelGGL(k_made_embed, dim3(blocks((uint64_t)B * H)), dim3(256), 0, S(stream), seq,
                       ld, B, L, q, H, W1t, b1, h1, static_cast<bf16*>(h1b));
    return launched();
}

static_assert(MAX_Q <= (uint32_t)EG_QMAX, "letter sums of the embedding gradient");

int mi3_made_embed_grad(void* stream, const uint8_t* seq, uint32_t ld, uint32_t B, uint32_t L,
                        uint32_t q, uint32_t H, const float* dh1, float* dW1t) {
    if (!seq || !dh1 || !dW1t || ld < B || q < 2 || q > MAX_Q)
        return mfail(MI3_EINVAL, "made_embed_grad: bad arguments");
    if (L == 0 || H == 0) return MI3_OK;
    hipLaunchKernelGGL(k_made_embed_grad, dim3(L * ((H + EG_T - 1) / EG_T)), dim3(EG_T), 0,
                       S(stream), seq, ld, B, L, q, H, dh1, dW1t);
    return launched();
}

int mi3_made_xent(void* stream, const float* logits, const uint8_t* seq, uint32_t ld, uint32_t B,
                  uint32_t L, uint32_t q, float gscale, float* nll, void* dlogits_bf16,
                  float* dlogits) {
    if (!logits || !seq || !nll || ld < B || q < 2 || q > MAX_Q || B == 0)
        return mfail(MI3_EINVAL, "made_xent: bad arguments");
    hipLaunchKernelGGL(k_made_xent, dim3(B), dim3(256), 0, S(stream), logits, seq, ld, B, L, q,
                       gscale, nll, static_cast<bf16*>(dlogits_bf16), dlogits);
    return launched();
}

int mi3_relu_grad(void* stream, uint64_t n, const float* dh, const float* hpre, float* out,
                  void* outb) {
    if (!dh || !hpre || (!out && !outb)) return mfail(MI3_EINVAL, "relu_grad: bad arguments");
    hipLaunchKernelGGL(k_relu_grad, dim3(blocks(n)), dim3(256), 0, S(stream), n, dh, hpre, out,
                       static_cast<bf16*>(outb));
    return launched();
}

int mi3_made_fused(void* stream, uint32_t L, uint32_t q, uint32_t H, uint32_t B, const float* W1t,
                   const float* b1, const void* W2b, const float* b2, const void* W3b,
                   const float* b3, uint64_t seed, uint32_t counter, uint32_t walker_base,
                   uint8_t* seq, uint32_t ld, float* nll, int draw) {
    if (!W1t || !b1 || !W2b || !b2 || !W3b || !b3 || !seq || !nll || L < 2 || q < 2 ||
        ld < B || !al16(W2b) || !al16(W3b))
        return mfail(MI3_EINVAL, "made_fused: bad arguments");
    if (q > (uint32_t)MF_QMAX || H == 0 || H % 32 || H > (uint32_t)MF_HMAX)
        return mfail(MI3_EUNSUP, "made_fused: needs q <= 25 and H a multiple of 32 up to 512");
    if (B == 0) return MI3_OK;
    hipLaunchKernelGGL(k_made_fused, dim3((B + MF_T - 1) / MF_T), dim3(MF_THREADS), 0, S(stream),
                       L, q, H, B, W1t, b1, static_cast<const bf16*>(W2b), b2,
                       static_cast<const bf16*>(W3b), b3, (uint32_t)seed, (uint32_t)(seed >> 32),
                       counter, walker_base, seq, ld, nll, draw);
    return launched();
}

int mi3_mh_accept(void* stream, uint32_t N, uint32_t L, uint8_t* cur, uint32_t ldc,
                  const uint8_t* prop, uint32_t ldp, float* Ecur, const float* Eprop,
                  float* nll_cur, const float* nll_prop, uint64_t seed, uint64_t counter,
                  uint32_t walker_base, uint32_t* naccept) {
    if (!cur || !prop || !Ecur || !Eprop || !nll_cur || !nll_prop || !naccept || ldc < N ||
        ldp < N)
        return mfail(MI3_EINVAL, "mh_accept: bad arguments");
    hipLaunchKernelGGL(k_mh_accept, dim3(blocks(N)), dim3(256), 0, S(stream), N, L, cur, ldc, prop,
                       ldp, Ecur, Eprop, nll_cur, nll_prop, (uint32_t)seed,
                       (uint32_t)(seed >> 32), counter, walker_base, naccept);
    return launched();
}

int mi3_adam_step(void* stream, uint64_t n, float* p, const float* g, float* m, float* v,
                  const float* mask, float lr, float beta1, float beta2, float eps, uint32_t t,
                  void* pb) {
    if (!p || !g || !m || !v || t == 0) return mfail(MI3_EINVAL, "adam: bad arguments");
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, S(stream), n, p, g, m, v, mask, lr,
                       beta1, beta2, eps, c1, c2, static_cast<bf16*>(pb));
    return launched();
}

int mi3_colsum(void* stream, uint32_t rows, uint32_t cols, const float* x, float* out) {
    if (!x || !out) return mfail(MI3_EINVAL, "colsum: bad arguments");
    if (cols == 0) return MI3_OK;
    hipLaunchKernelGGL(k_colsum, dim3((cols + CS_COLS - 1) / CS_COLS), dim3(CS_COLS * CS_ROWL), 0,
                       S(stream), rows, cols, x, out);
    return launched();
}

int mi3_to_bf16(void* stream, uint64_t n, const float* p, void* pb) {
    if (!p || !pb) return mfail(MI3_EINVAL, "to_bf16: bad arguments");
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, S(stream), n, p,
                       static_cast<bf16*>(pb));
    return launched();
}

}  // extern "C"
