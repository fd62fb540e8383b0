/*
 * C restatement of the Mi3 inference-round kernels (test infrastructure).
 *
 * Checker and CPU port for the HIP engine; never part of the product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  Same mathematics as oracle/potts_oracle.py, in C with OpenMP
 * over walkers/sequences so that checks at production sizes finish quickly.
 *
 *   orc_philox4x32_10   Philox4x32-10 (Salmon et al., SC'11)
 *   orc_positions       shared site sequence (replaces mcmcGPU.py:447-458)
 *   orc_metropolis      single-site Metropolis, float64 dE, acceptance margins
 *                       (mcmc.cl:360-458; extras/mcmcCPUgenThreaded.c:68-84)
 *   orc_metropolis_f32  the same chain in float32 arithmetic, as the reference
 *                       CPU sampler computes it (mcmcCPUgenThreaded.c:68-84);
 *                       this is the timed CPU port
 *   orc_energies        float64 Potts energies (getSeqEnergies.py:38-60)
 *   orc_bicounts        pair counts (getMarginals.py:55-87)
 *   orc_weighted_marg   float64 weighted pair histograms (mcmc.cl:646-718)
 *   orc_gen_indep       independent-site reseed (mcmc.cl:170-244)
 *
 * Sequences are row-major (N, L) uint8; couplings are the packed
 * (L(L-1)/2, q*q) float32 format (UserGuide.md:176-180).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TAG_WALK 0x57414C4Bu
#define TAG_POS 0x504F5349u
#define TAG_INDEP 0x494E4450u

static inline void philox(const uint32_t c_in[4], uint32_t k0, uint32_t k1,
                          uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key[0], key[1], out);
}

static inline uint32_t mulhi32(uint32_t x, uint32_t n) {
    return (uint32_t)(((uint64_t)x * n) >> 32);
}

static inline float uniform24(uint32_t y) {
    return (float)(y >> 8) * 0x1.0p-24f;
}

void orc_positions(uint64_t seed, uint32_t L, uint64_t t0, uint32_t n, uint32_t *pos) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (uint32_t t = 0; t < n; t++) {
        uint64_t g = t0 + t;
        uint32_t c[4] = {(uint32_t)g, (uint32_t)(g >> 32), 0u, TAG_POS}, r[4];
        philox(c, k0, k1, r);
        pos[t] = mulhi32(r[0], L);
    }
}

static inline void walker_draw(uint32_t k0, uint32_t k1, uint32_t walker, uint64_t t,
                               uint32_t q, uint32_t *b, float *u) {
    uint64_t h = t >> 1;
    uint32_t c[4] = {(uint32_t)h, (uint32_t)(h >> 32), walker, TAG_WALK}, r[4];
    philox(c, k0, k1, r);
    uint32_t x = (t & 1) ? r[2] : r[0];
    uint32_t y = (t & 1) ? r[3] : r[1];
    *b = mulhi32(x, q);
    *u = uniform24(y);
}

/* unpack to the CPU sampler's (L*L, q*q) layout, coupling(i,j,a,b) =
 * full[(L*i + j)*q*q + a*q + b] (mcmcCPUgenThreaded.c:35, 193-206) */
static float *unpack_f32(const float *J, uint32_t L, uint32_t q) {
    size_t qq = (size_t)q * q;
    float *F = (float *)calloc((size_t)L * L * qq, sizeof(float));
    size_t n = 0;
    for (uint32_t i = 0; i + 1 < L; i++)
        for (uint32_t j = i + 1; j < L; j++, n++)
            for (uint32_t a = 0; a < q; a++)
                for (uint32_t b = 0; b < q; b++) {
                    float v = J[n * qq + a * q + b];
                    F[((size_t)L * i + j) * qq + a * q + b] = v;
                    F[((size_t)L * j + i) * qq + b * q + a] = v;
                }
    return F;
}

/* float64-dE chain with acceptance margins (the checker). */
int orc_metropolis(const float *J, uint32_t L, uint32_t q, uint8_t *seqs, uint32_t N,
                   uint64_t seed, uint64_t t0, uint32_t nsteps, float beta,
                   uint32_t walker0, double *margin, int nthreads) {
    float *F = unpack_f32(J, L, q);
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * (nsteps ? nsteps : 1));
    if (!F || !pos) return -1;
    orc_positions(seed, L, t0, nsteps, pos);
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    size_t qq = (size_t)q * q;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long w = 0; w < (long)N; w++) {
        uint8_t *s = seqs + (size_t)w * L;
        double mg = INFINITY;
        for (uint32_t t = 0; t < nsteps; t++) {
            uint32_t p = pos[t], b;
            float u;
            walker_draw(k0, k1, walker0 + (uint32_t)w, t0 + t, q, &b, &u);
            uint32_t a = s[p];
            const float *row = F + (size_t)L * p * qq;
            double dE = 0;
            for (uint32_t m = 0; m < L; m++) {
                if (m == p) continue;
                dE += (double)row[m * qq + b * q + s[m]] - (double)row[m * qq + a * q + s[m]];
            }
            double pr = exp(-(double)beta * dE);
            double d = fabs(pr - (double)u);
            if (d < mg) mg = d;
            if (pr > (double)u) s[p] = (uint8_t)b;
        }
        if (margin) margin[w] = mg;
    }
    free(F);
    free(pos);
    return 0;
}

/* float32 chain: the arithmetic of the reference CPU sampler's step()
 * (mcmcCPUgenThreaded.c:68-84), with this engine's random streams. */
int orc_metropolis_f32(const float *J, uint32_t L, uint32_t q, uint8_t *seqs, uint32_t N,
                       uint64_t seed, uint64_t t0, uint32_t nsteps, float beta,
                       uint32_t walker0, int nthreads) {
    float *F = unpack_f32(J, L, q);
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * (nsteps ? nsteps : 1));
    if (!F || !pos) return -1;
    orc_positions(seed, L, t0, nsteps, pos);
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    size_t qq = (size_t)q * q;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long w = 0; w < (long)N; w++) {
        uint8_t *s = seqs + (size_t)w * L;
        for (uint32_t t = 0; t < nsteps; t++) {
            uint32_t p = pos[t], b;
            float u;
            walker_draw(k0, k1, walker0 + (uint32_t)w, t0 + t, q, &b, &u);
            uint32_t a = s[p];
            const float *row = F + (size_t)L * p * qq;
            float dE = 0;
            for (uint32_t m = 0; m < L; m++)
                dE += row[m * qq + b * q + s[m]] - row[m * qq + a * q + s[m]];
            if (expf(-beta * dE) > u) s[p] = (uint8_t)b;
        }
    }
    free(F);
    free(pos);
    return 0;
}

int orc_energies(const float *J, uint32_t L, uint32_t q, const uint8_t *seqs, uint32_t N,
                 double *E, int nthreads) {
    size_t qq = (size_t)q * q;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long w = 0; w < (long)N; w++) {
        const uint8_t *s = seqs + (size_t)w * L;
        double e = 0;
        size_t n = 0;
        for (uint32_t i = 0; i + 1 < L; i++)
            for (uint32_t j = i + 1; j < L; j++, n++)
                e += (double)J[n * qq + q * s[i] + s[j]];
        E[w] = e;
    }
    return 0;
}

/* site-major copy (L rows of N residues), so that each pair's scan reads two
 * contiguous rows */
static uint8_t *site_major(const uint8_t *seqs, uint32_t N, uint32_t L) {
    uint8_t *T = (uint8_t *)malloc((size_t)N * L + 1);
    if (!T) return NULL;
    for (uint32_t w = 0; w < N; w++)
        for (uint32_t i = 0; i < L; i++) T[(size_t)i * N + w] = seqs[(size_t)w * L + i];
    return T;
}

int orc_bicounts(const uint8_t *seqs, uint32_t N, uint32_t L, uint32_t q, uint32_t *counts,
                 int nthreads) {
    size_t qq = (size_t)q * q;
    size_t nP = (size_t)L * (L - 1) / 2;
    memset(counts, 0, nP * qq * sizeof(uint32_t));
    uint8_t *T = site_major(seqs, N, L);
    if (!T) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long i = 0; i < (long)L - 1; i++) {
        size_t n0 = (size_t)i * (2 * L - i - 1) / 2;
        const uint8_t *si = T + (size_t)i * N;
        for (uint32_t j = (uint32_t)i + 1; j < L; j++) {
            uint32_t *c = counts + (n0 + (j - i - 1)) * qq;
            const uint8_t *sj = T + (size_t)j * N;
            for (uint32_t w = 0; w < N; w++) c[q * si[w] + sj[w]]++;
        }
    }
    free(T);
    return 0;
}

int orc_weighted_marg(const uint8_t *seqs, const float *wts, uint32_t N, uint32_t L,
                      uint32_t q, double *F, int nthreads) {
    size_t qq = (size_t)q * q;
    size_t nP = (size_t)L * (L - 1) / 2;
    memset(F, 0, nP * qq * sizeof(double));
    uint8_t *T = site_major(seqs, N, L);
    if (!T) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long i = 0; i < (long)L - 1; i++) {
        size_t n0 = (size_t)i * (2 * L - i - 1) / 2;
        const uint8_t *si = T + (size_t)i * N;
        for (uint32_t j = (uint32_t)i + 1; j < L; j++) {
            double *c = F + (n0 + (j - i - 1)) * qq;
            const uint8_t *sj = T + (size_t)j * N;
            for (uint32_t w = 0; w < N; w++) c[q * si[w] + sj[w]] += (double)wts[w];
        }
    }
    free(T);
    return 0;
}

int orc_gen_indep(const float *cprob, uint32_t L, uint32_t q, uint64_t seed, uint32_t call,
                  uint32_t N, uint32_t walker0, uint8_t *seqs) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (uint32_t w = 0; w < N; w++) {
        for (uint32_t k = 0; 4 * k < L; k++) {
            uint32_t c[4] = {k, call, walker0 + w, TAG_INDEP}, r[4];
            philox(c, k0, k1, r);
            for (uint32_t e = 0; e < 4 && 4 * k + e < L; e++) {
                uint32_t i = 4 * k + e;
                float u = uniform24(r[e]);
                uint32_t res = 0;
                for (uint32_t a = 0; a + 1 < q; a++)
                    if (u >= cprob[(size_t)i * (q - 1) + a]) res++;
                seqs[(size_t)w * L + i] = (uint8_t)res;
            }
        }
    }
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
