// Host-side launchers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi3 {

struct SweepArgs {
    const float* Ju;  // chunked layout, chunk stride chunk_floats(q)
    uint8_t* seq;
    const uint32_t* pos;
    uint32_t ld, nwalk, L, q, nchunk, nsteps, k0, k1, walker_base;
    uint64_t step0;
    float nbeta;
    uint32_t wpb;  // walkers per work-group: 128, 256 or 512 (0 = default)
};

struct EnergyArgs {
    const float* Ju;
    const uint8_t* seq;
    float* E;
    uint32_t ld, nseq, L, q, nchunk;
};

// Walkers per work-group of the sweep kernel: 128, 256 or 512 (a request of
// 512 runs tiles of 256 for L > 384 and q <= 8); the walker count must be a
// multiple of 512.
constexpr uint32_t SWEEP_WALKERS_PER_BLOCK = 512;
constexpr uint32_t SWEEP_DEFAULT_WPB = 512;

// Largest q / L the register-resident kernels are instantiated for.
constexpr uint32_t MAX_Q = 25;
constexpr uint32_t MAX_L = 1024;

// Threads per work-group of the sweep / energy kernels.
constexpr uint32_t POTTS_THREADS = 512;

// float4 loads per thread to stage one chunk of q*q*32 floats; the chunked
// coupling layout pads every chunk to POTTS_THREADS * pf_class(q) float4 so
// that staging needs no bounds checks.
inline uint32_t pf_class(uint32_t q) {
    uint32_t need = (q * q * 32 / 4 + POTTS_THREADS - 1) / POTTS_THREADS;
    if (need <= 1) return 1;
    if (need <= 4) return 4;
    if (need <= 7) return 7;
    return 10;
}

// Floats per (padded) chunk.
inline uint32_t chunk_floats(uint32_t q) { return pf_class(q) * POTTS_THREADS * 4; }

// The chunked couplings of an L-site model are addressed with 32-bit byte
// offsets (buffer resources): true when they fit.
inline bool chunked_fits(uint32_t L, uint32_t q) {
    const uint64_t nchunk = (L + 31) / 32;
    return (uint64_t)L * nchunk * chunk_floats(q) * 4 < (1ull << 32);
}

// One model of a batch, as the persistent sweep reads it (device memory).
struct BatchModel {
    const float* Ju;        // chunked couplings
    uint8_t* seq;           // site-major walkers, L rows of ld bytes
    const uint32_t* pos;    // site list of this launch
    const uint64_t* step;   // global step counter (even)
    uint32_t ld, L, q, nchunk, nsteps, k0, k1, walker_base;
    float nbeta;
    uint32_t pad_;
};

// A work-group tile of a batch: walkers [block*W, (block+1)*W) of `model`.
struct BatchTile {
    uint32_t model, block;
};

// Returns a hipError_t; grid/LDS sizing done inside.
hipError_t launch_sweep(const SweepArgs& a, hipStream_t st);
hipError_t launch_energies(const EnergyArgs& a, hipStream_t st);

// The sweep's coupling-row staging alone (k_stage_probe) for the sites of
// a.pos, one work-group per 512 walkers; inflight 1 or 2 chunk copies.
hipError_t launch_stage_probe(const SweepArgs& a, int inflight, hipStream_t st);

// State words per lane (4 residues each) for L sites: ceil(L / 128).
inline uint32_t state_words(uint32_t L) { return (L + 127) / 128; }

// Persistent sweep over `ntiles` tiles of models that share state_words(L)
// == kw and pf_class(q) == pf, with `wpb` walkers per tile (128, 256 or 512),
// on `grid` work-groups.  `counter` (one u32) must be zero at launch.
hipError_t launch_sweep_batch(uint32_t kw, uint32_t pf, uint32_t wpb, const BatchModel* models,
                              const BatchTile* tiles, uint32_t ntiles, uint32_t* counter,
                              uint32_t grid, hipStream_t st);

// Chunked symmetric coupling layout from packed couplings (k_unpack_chunked).
hipError_t launch_unpack_chunked(const float* J, float* out, uint32_t L, uint32_t q,
                                 uint32_t nchunk, hipStream_t st);

}  // namespace mi3
