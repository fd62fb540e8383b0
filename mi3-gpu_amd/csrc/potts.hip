// Instantiation and launch of the register-resident sweep / energy kernels.
#include "k_potts.hip"
#include "launch.h"

namespace mi3 {

namespace {
constexpr int NWAVES = 8;
constexpr int NT = 64 * NWAVES;

// Launch shapes (walker slots per half-wave S, waves NW): W = 2*S*NW walkers
// per work-group.  16 waves of at most 128 registers (4 waves per SIMD) hide
// the gathers' latency better than 8 waves of 256, wherever the state fits:
// measured on MI355X at L=120/232/300/450, q=21 and q=4.
// sequence slots per half-wave of the energy kernel: every staged coupling
// chunk is read once per sequence, so more sequences per work-group cut the
// L2 -> LDS traffic per sequence.  16 slots in 16 waves (512 sequences per
// work-group) up to four state words, 8 beyond: with byte gather addresses
// this beats 32 slots in 8 waves for two state words (2.15 against 2.45 ms
// at L=232 q=21, 1.31 against 1.67 ms at L=200 q=8).  One state word with the
// single-pass chunks of q <= 8 takes 8 slots: 45 registers, so two
// work-groups share a CU and one's barriers overlap the other's gathers
// (0.33 against 0.37 ms at L=93 q=4, 0.35 against 0.38 at L=93 q=8); with
// larger chunks the LDS holds one work-group and 8 slots lose (0.73 against
// 0.49 ms at L=93 q=21)
template <int KW, int PF>
constexpr int eslots() { return (KW == 1 && PF == 1) ? 8 : KW <= 4 ? 16 : 8; }
template <int KW>
constexpr int ewaves() { return 16; }

static_assert(NT == POTTS_THREADS, "thread count");

template <typename F>
hipError_t set_lds(F kernel, size_t bytes) {
    static_assert(sizeof(F) > 0, "");
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int KW, int PF, int S, int NW = NWAVES>
hipError_t sweep_kws(const SweepArgs& a, hipStream_t st) {
    constexpr uint32_t W = 2 * S * NW;
    auto kern = k_sweep<KW, S, NW, PF>;
    size_t lds = 2ull * chunk_floats(a.q) * sizeof(float);
    hipError_t e = set_lds(kern, lds);
    if (e != hipSuccess) return e;
    if (a.nwalk % W) return hipErrorInvalidValue;
    dim3 grid(a.nwalk / W);
    hipLaunchKernelGGL(kern, grid, dim3(64 * NW), lds, st, a.Ju, a.seq, a.pos, a.ld, a.L, a.q,
                       a.nchunk, a.nsteps, a.k0, a.k1, a.step0, a.walker_base, a.nbeta);
    return hipGetLastError();
}

template <int KW, int PF>
hipError_t energies_kw(const EnergyArgs& a, hipStream_t st) {
    constexpr int NW = ewaves<KW>();
    constexpr uint32_t W = 2 * eslots<KW, PF>() * NW;
    // with four state words (L = 385..512), the 57 KB chunks of q = 17..21
    // stream through the overlapping four-slot ring (all 160 KB of LDS):
    // 7.10 against 7.94 ms at L=400, 10.7 against 11.8 ms at L=500.  The
    // others are double-buffered (the ring took 6% longer at L=300 and 1% at
    // L=232)
    constexpr bool OVL = PF == 7 && KW == 4;
    auto kern = k_energies<KW, eslots<KW, PF>(), NW, PF, OVL>;
    size_t lds = OVL ? (size_t)OverlapRing::BYTES : 2ull * chunk_floats(a.q) * sizeof(float);
    hipError_t e = set_lds(kern, lds);
    if (e != hipSuccess) return e;
    dim3 grid((a.nseq + W - 1) / W);
    hipLaunchKernelGGL(kern, grid, dim3(64 * NW), lds, st, a.Ju, a.seq, a.E, a.ld, a.nseq, a.L,
                       a.q, a.nchunk);
    return hipGetLastError();
}

template <template <int, int> class Disp, typename A>
hipError_t dispatch(const A& a, hipStream_t st, uint32_t nchunk, uint32_t q) {
    uint32_t kw = (nchunk + 3) / 4;
    int pf = (int)pf_class(q);
#define MI3_CASE(KW_, PF_) \
    if (kw == KW_ && pf == PF_) return Disp<KW_, PF_>::run(a, st);
    MI3_CASE(1, 1) MI3_CASE(1, 4) MI3_CASE(1, 7) MI3_CASE(1, 10)
    MI3_CASE(2, 1) MI3_CASE(2, 4) MI3_CASE(2, 7) MI3_CASE(2, 10)
    MI3_CASE(3, 1) MI3_CASE(3, 4) MI3_CASE(3, 7) MI3_CASE(3, 10)
    MI3_CASE(4, 1) MI3_CASE(4, 4) MI3_CASE(4, 7) MI3_CASE(4, 10)
    MI3_CASE(5, 1) MI3_CASE(5, 4) MI3_CASE(5, 7) MI3_CASE(5, 10)
    MI3_CASE(6, 1) MI3_CASE(6, 4) MI3_CASE(6, 7) MI3_CASE(6, 10)
    MI3_CASE(7, 1) MI3_CASE(7, 4) MI3_CASE(7, 7) MI3_CASE(7, 10)
    MI3_CASE(8, 1) MI3_CASE(8, 4) MI3_CASE(8, 7) MI3_CASE(8, 10)
#undef MI3_CASE
    return hipErrorInvalidValue;
}

template <int KW, int PF>
struct SweepDisp {
    static hipError_t run(const SweepArgs& a, hipStream_t st) {
        uint32_t w = a.wpb ? a.wpb : SWEEP_DEFAULT_WPB;
        if constexpr (KW >= 5) {
            // five to eight state words: 8 slots at most fit in 128 registers
            if (w <= 128) return sweep_kws<KW, PF, 4, 16>(a, st);
            return sweep_kws<KW, PF, 8, 16>(a, st);
        } else {
            if (w <= 128) return sweep_kws<KW, PF, 8, 8>(a, st);
            if (w <= 256) return sweep_kws<KW, PF, 8, 16>(a, st);
            // four state words do not fit 16 slots in 128 registers; the
            // spills still pay for q > 8, not for the small chunks of q <= 8
            if constexpr (KW == 4 && PF == 1)
                return sweep_kws<KW, PF, 8, 16>(a, st);
            else
                return sweep_kws<KW, PF, 16, 16>(a, st);
        }
    }
};
template <int KW, int PF>
struct EnergyDisp {
    static hipError_t run(const EnergyArgs& a, hipStream_t st) { return energies_kw<KW, PF>(a, st); }
};
}  // namespace

hipError_t launch_sweep(const SweepArgs& a, hipStream_t st) {
    if (a.nwalk % SWEEP_WALKERS_PER_BLOCK || a.q > MAX_Q || a.L > MAX_L || a.L < 2)
        return hipErrorInvalidValue;
    return dispatch<SweepDisp>(a, st, a.nchunk, a.q);
}

namespace {
template <int KW, int PF, int S, int NW = NWAVES>
hipError_t batch_kws(const BatchModel* models, const BatchTile* tiles, uint32_t ntiles,
                     uint32_t* counter, uint32_t grid, hipStream_t st) {
    auto kern = k_sweep_batch<KW, S, NW, PF>;
    size_t lds = 2ull * PF * POTTS_THREADS * 4 * sizeof(float);  // two padded chunks
    hipError_t e = set_lds(kern, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * NW), lds, st, models, tiles, ntiles, counter);
    return hipGetLastError();
}

template <int KW, int PF>
hipError_t batch_kw(uint32_t wpb, const BatchModel* models, const BatchTile* tiles,
                    uint32_t ntiles, uint32_t* counter, uint32_t grid, hipStream_t st) {
    if constexpr (KW >= 5) {
        if (wpb == 128) return batch_kws<KW, PF, 4, 16>(models, tiles, ntiles, counter, grid, st);
        if (wpb == 256) return batch_kws<KW, PF, 8, 16>(models, tiles, ntiles, counter, grid, st);
        return hipErrorInvalidValue;
    } else {
        if (wpb == 128) return batch_kws<KW, PF, 8, 8>(models, tiles, ntiles, counter, grid, st);
        if (wpb == 256) return batch_kws<KW, PF, 8, 16>(models, tiles, ntiles, counter, grid, st);
        if (wpb == 512) return batch_kws<KW, PF, 16, 16>(models, tiles, ntiles, counter, grid, st);
        return hipErrorInvalidValue;
    }
}
}  // namespace

hipError_t launch_sweep_batch(uint32_t kw, uint32_t pf, uint32_t wpb, const BatchModel* models,
                              const BatchTile* tiles, uint32_t ntiles, uint32_t* counter,
                              uint32_t grid, hipStream_t st) {
    if (grid == 0 || ntiles == 0) return hipSuccess;
#define MI3_BCASE(KW_, PF_) \
    if (kw == KW_ && pf == PF_) return batch_kw<KW_, PF_>(wpb, models, tiles, ntiles, counter, grid, st);
    MI3_BCASE(1, 1) MI3_BCASE(1, 4) MI3_BCASE(1, 7) MI3_BCASE(1, 10)
    MI3_BCASE(2, 1) MI3_BCASE(2, 4) MI3_BCASE(2, 7) MI3_BCASE(2, 10)
    MI3_BCASE(3, 1) MI3_BCASE(3, 4) MI3_BCASE(3, 7) MI3_BCASE(3, 10)
    MI3_BCASE(4, 1) MI3_BCASE(4, 4) MI3_BCASE(4, 7) MI3_BCASE(4, 10)
    MI3_BCASE(5, 1) MI3_BCASE(5, 4) MI3_BCASE(5, 7) MI3_BCASE(5, 10)
    MI3_BCASE(6, 1) MI3_BCASE(6, 4) MI3_BCASE(6, 7) MI3_BCASE(6, 10)
    MI3_BCASE(7, 1) MI3_BCASE(7, 4) MI3_BCASE(7, 7) MI3_BCASE(7, 10)
    MI3_BCASE(8, 1) MI3_BCASE(8, 4) MI3_BCASE(8, 7) MI3_BCASE(8, 10)
#undef MI3_BCASE
    return hipErrorInvalidValue;
}

namespace {
template <int PF, int INF>
hipError_t stage_probe_pf(const SweepArgs& a, hipStream_t st) {
    constexpr int NW = 16;
    auto kern = k_stage_probe<PF, NW, INF>;
    size_t lds = 2ull * chunk_floats(a.q) * sizeof(float);
    hipError_t e = set_lds(kern, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(a.nwalk / 512), dim3(64 * NW), lds, st, a.Ju, a.pos, a.L,
                       a.nchunk, a.nsteps);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_stage_probe(const SweepArgs& a, int inflight, hipStream_t st) {
    if (a.nwalk % 512 || a.q > MAX_Q || a.L > MAX_L || a.L < 2) return hipErrorInvalidValue;
    const uint32_t pf = pf_class(a.q);
#define MI3_PCASE(PF_)                                                             \
    if (pf == PF_) return inflight == 2 ? stage_probe_pf<PF_, 2>(a, st)             \
                                        : stage_probe_pf<PF_, 1>(a, st);
    MI3_PCASE(1) MI3_PCASE(4) MI3_PCASE(7) MI3_PCASE(10)
#undef MI3_PCASE
    return hipErrorInvalidValue;
}

hipError_t launch_energies(const EnergyArgs& a, hipStream_t st) {
    if (a.q > MAX_Q || a.L > MAX_L || a.L < 2) return hipErrorInvalidValue;
    if (a.nseq == 0) return hipSuccess;
    return dispatch<EnergyDisp>(a, st, a.nchunk, a.q);
}

}  // namespace mi3
