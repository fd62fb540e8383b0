// libmi3hip.so, batched models: many independent Potts models of uneven L
// and q, each with its own walkers, couplings and random streams, swept by
// one persistent kernel per (state words, chunk class) class.  One round
// (positions of every model, the class sweeps on parallel branches, the
// step-counter advance) is captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "common.h"
#include "launch.h"
#include "mi3hip.h"

using namespace mi3;

namespace {

// Site lists of every model for this launch:
// pos_m[t] = mulhi(Philox((step_m + t, 0, TAG_POS)).x, L_m), the stream of
// a single-model context keyed by the model's seed.
__global__ void k_batch_positions(const BatchModel* __restrict__ models, uint32_t* const* pos) {
    const BatchModel m = models[blockIdx.y];
    const uint64_t step0 = *m.step;
    uint32_t* p = pos[blockIdx.y];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < m.nsteps;
         t += gridDim.x * blockDim.x) {
        const uint64_t g = step0 + t;
        uint4 r = philox4x32_10(make_uint4((uint32_t)g, (uint32_t)(g >> 32), 0u, TAG_POS), m.k0,
                                m.k1);
        p[t] = mulhi_range(r.x, m.L);
    }
}

// Step counters move past this launch, kept even (as a single-model context).
__global__ void k_batch_advance(const BatchModel* __restrict__ models, uint64_t* steps,
                                uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) steps[i] += models[i].nsteps + (models[i].nsteps & 1u);
}

__global__ void k_batch_zero(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

struct ClassKey {
    uint32_t kw, pf;
    bool operator<(const ClassKey& o) const { return kw != o.kw ? kw < o.kw : pf < o.pf; }
};

struct ClassRun {
    std::vector<uint32_t> models;
    uint32_t wpb = 0;        // chosen walkers per tile
    BatchTile* tiles = nullptr;
    uint32_t ntiles = 0;
    hipStream_t side = nullptr;
};

}  // namespace

struct mi3_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint64_t seed = 0;
    uint32_t n = 0;
    std::vector<uint32_t> L, q, N, nchunk;
    std::vector<float*> Jp, Ju;      // packed / chunked couplings
    std::vector<uint8_t*> seq;
    std::vector<uint32_t*> pos;
    std::vector<uint32_t> pos_cap;
    std::vector<BatchModel> host_models;
    BatchModel* models = nullptr;    // device copy
    uint32_t** pos_ptrs = nullptr;   // device array of the pos pointers
    uint64_t* steps = nullptr;
    uint32_t* counters = nullptr;    // one per class
    std::map<ClassKey, ClassRun> classes;
    uint32_t nsweeps = 0;            // of the captured graph
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t ncu = 0;
};

namespace {

int bfail(mi3_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

#define BCHK(b, call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bfail(b, MI3_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

thread_local std::string g_batch_create_error;

void drop_graph(mi3_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->nsweeps = 0;
}

// Walkers per tile allowed for a class.
std::vector<uint32_t> tile_options(uint32_t kw) {
    if (kw >= 5) return {128, 256};
    return {128, 256, 512};
}

// Tiles of a class for a walkers-per-tile choice; larger models first so
// that the tail of the persistent loop is short.
int build_tiles(mi3_batch* b, ClassRun& cr, uint32_t wpb) {
    std::vector<BatchTile> t;
    std::vector<uint32_t> order = cr.models;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return b->L[x] > b->L[y]; });
    for (uint32_t m : order)
        for (uint32_t blk = 0; blk < b->N[m] / wpb; blk++) t.push_back({m, blk});
    if (cr.tiles) hipFree(cr.tiles);
    cr.tiles = nullptr;
    BCHK(b, hipMalloc(&cr.tiles, std::max<size_t>(1, t.size()) * sizeof(BatchTile)));
    BCHK(b, hipMemcpy(cr.tiles, t.data(), t.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
    cr.ntiles = (uint32_t)t.size();
    cr.wpb = wpb;
    return MI3_OK;
}

uint32_t class_grid(mi3_batch* b, const ClassRun& cr) {
    return std::min<uint32_t>(cr.ntiles, b->ncu ? b->ncu : 256);
}

// Per-model descriptors for nsweeps sweeps (nsteps = nsweeps * L).
int upload_models(mi3_batch* b, uint32_t nsweeps) {
    for (uint32_t m = 0; m < b->n; m++) {
        const uint32_t ns = nsweeps * b->L[m];
        if (b->pos_cap[m] < ns) {
            if (b->pos[m]) hipFree(b->pos[m]);
            b->pos[m] = nullptr;
            BCHK(b, hipMalloc(&b->pos[m], (size_t)ns * 4));
            b->pos_cap[m] = ns;
        }
        BatchModel& d = b->host_models[m];
        d.Ju = b->Ju[m];
        d.seq = b->seq[m];
        d.pos = b->pos[m];
        d.step = b->steps + m;
        d.ld = b->N[m];
        d.L = b->L[m];
        d.q = b->q[m];
        d.nchunk = b->nchunk[m];
        d.nsteps = ns;
        d.k0 = (uint32_t)(b->seed + m);
        d.k1 = (uint32_t)((b->seed + m) >> 32);
        d.walker_base = 0;
        d.nbeta = -1.f;
        d.pad_ = 0;
    }
    BCHK(b, hipMemcpy(b->models, b->host_models.data(), b->n * sizeof(BatchModel),
                      hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->pos_ptrs, b->pos.data(), b->n * sizeof(uint32_t*),
                      hipMemcpyHostToDevice));
    return MI3_OK;
}

uint32_t max_steps(mi3_batch* b) {
    uint32_t mx = 0;
    for (auto& d : b->host_models) mx = std::max(mx, d.nsteps);
    return mx;
}

// Enqueue one round on b->stream (also used under stream capture):
// positions, class sweeps forked onto side streams, join, step advance.
int enqueue_round(mi3_batch* b, std::vector<hipEvent_t>& events) {
    const uint32_t nclass = (uint32_t)b->classes.size();
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, b->stream, b->counters, nclass);
    const uint32_t mx = max_steps(b);
    hipLaunchKernelGGL(k_batch_positions, dim3(std::max(1u, std::min(64u, (mx + 255) / 256)), b->n),
                       dim3(256), 0, b->stream, (const BatchModel*)b->models,
                       (uint32_t* const*)b->pos_ptrs);
    BCHK(b, hipGetLastError());
    hipEvent_t fork;
    BCHK(b, hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    events.push_back(fork);
    BCHK(b, hipEventRecord(fork, b->stream));
    uint32_t ci = 0;
    int rc = MI3_OK;
    for (auto& kv : b->classes) {
        ClassRun& cr = kv.second;
        hipStream_t s = cr.side;
        if (hipStreamWaitEvent(s, fork, 0) != hipSuccess) { rc = bfail(b, MI3_EHIP, "fork"); break; }
        hipError_t e = launch_sweep_batch(kv.first.kw, kv.first.pf, cr.wpb, b->models, cr.tiles,
                                          cr.ntiles, b->counters + ci, class_grid(b, cr), s);
        if (e != hipSuccess) {
            rc = bfail(b, MI3_EHIP, std::string("batch sweep: ") + hipGetErrorString(e));
            break;
        }
        hipEvent_t j;
        if (hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
            rc = bfail(b, MI3_EHIP, "event");
            break;
        }
        events.push_back(j);
        if (hipEventRecord(j, s) != hipSuccess || hipStreamWaitEvent(b->stream, j, 0) != hipSuccess) {
            rc = bfail(b, MI3_EHIP, "join");
            break;
        }
        ci++;
    }
    if (rc == MI3_OK) {
        hipLaunchKernelGGL(k_batch_advance, dim3((b->n + 255) / 256), dim3(256), 0, b->stream,
                           (const BatchModel*)b->models, b->steps, b->n);
        if (hipGetLastError() != hipSuccess) rc = bfail(b, MI3_EHIP, "advance");
    }
    return rc;
}

}  // namespace

extern "C" {

int mi3_batch_create(int device, uint32_t nmodels, const uint32_t* L, const uint32_t* q,
                     const uint32_t* nwalkers, uint64_t seed, mi3_batch** out) {
    if (!out) return MI3_EINVAL;
    *out = nullptr;
    if (nmodels == 0 || !L || !q || !nwalkers) {
        g_batch_create_error = "need at least one model";
        return MI3_EINVAL;
    }
    for (uint32_t m = 0; m < nmodels; m++) {
        if (L[m] < 2 || q[m] < 2 || q[m] > MAX_Q || L[m] > MAX_L || !chunked_fits(L[m], q[m])) {
            g_batch_create_error = "model " + std::to_string(m) + ": need 2 <= L <= " +
                                   std::to_string(MAX_L) + " and 2 <= q <= " +
                                   std::to_string(MAX_Q) +
                                   ", with chunked couplings under 4 GiB";
            return MI3_EUNSUP;
        }
        if (nwalkers[m] == 0 || nwalkers[m] % 512) {
            g_batch_create_error = "model " + std::to_string(m) +
                                   ": walkers must be a positive multiple of 512";
            return MI3_EINVAL;
        }
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_batch_create_error = "bad device";
        return MI3_EHIP;
    }
    mi3_batch* b = new mi3_batch();
    b->dev = device;
    b->seed = seed;
    b->n = nmodels;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) b->ncu = prop.multiProcessorCount;
    int rc = MI3_OK;
    auto bad = [&](const char* what) {
        g_batch_create_error = what;
        rc = MI3_ENOMEM;
    };
    if (hipStreamCreate(&b->stream) != hipSuccess) bad("stream");
    b->L.assign(L, L + nmodels);
    b->q.assign(q, q + nmodels);
    b->N.assign(nwalkers, nwalkers + nmodels);
    b->nchunk.resize(nmodels);
    b->Jp.assign(nmodels, nullptr);
    b->Ju.assign(nmodels, nullptr);
    b->seq.assign(nmodels, nullptr);
    b->pos.assign(nmodels, nullptr);
    b->pos_cap.assign(nmodels, 0);
    b->host_models.resize(nmodels);
    for (uint32_t m = 0; m < nmodels && rc == MI3_OK; m++) {
        const uint32_t nc = (L[m] + CH - 1) / CH;
        b->nchunk[m] = nc;
        const size_t np = (size_t)L[m] * (L[m] - 1) / 2 * q[m] * q[m];
        if (hipMalloc(&b->Jp[m], np * 4) != hipSuccess ||
            hipMalloc(&b->Ju[m], (size_t)L[m] * nc * chunk_floats(q[m]) * 4) != hipSuccess ||
            hipMalloc(&b->seq[m], (size_t)L[m] * nwalkers[m]) != hipSuccess)
            bad("device memory for a model");
        else {
            hipMemset(b->Jp[m], 0, np * 4);
            hipMemset(b->seq[m], 0, (size_t)L[m] * nwalkers[m]);
        }
        ClassKey k{state_words(L[m]), pf_class(q[m])};
        b->classes[k].models.push_back(m);
    }
    if (rc == MI3_OK &&
        (hipMalloc(&b->models, nmodels * sizeof(BatchModel)) != hipSuccess ||
         hipMalloc(&b->pos_ptrs, nmodels * sizeof(uint32_t*)) != hipSuccess ||
         hipMalloc(&b->steps, nmodels * sizeof(uint64_t)) != hipSuccess ||
         hipMalloc(&b->counters, b->classes.size() * sizeof(uint32_t)) != hipSuccess))
        bad("device memory for descriptors");
    if (rc == MI3_OK) {
        hipMemset(b->steps, 0, nmodels * sizeof(uint64_t));
        for (uint32_t m = 0; m < nmodels && rc == MI3_OK; m++)
            if (launch_unpack_chunked(b->Jp[m], b->Ju[m], L[m], q[m], b->nchunk[m], b->stream) !=
                hipSuccess)
                bad("unpack");
        for (auto& kv : b->classes) {
            if (rc != MI3_OK) break;
            if (hipStreamCreate(&kv.second.side) != hipSuccess) { bad("stream"); break; }
            // default tile: the largest allowed
            if (build_tiles(b, kv.second, tile_options(kv.first.kw).back()) != MI3_OK)
                bad("tiles");
        }
        if (hipStreamSynchronize(b->stream) != hipSuccess) bad("sync");
    }
    if (rc != MI3_OK) {
        mi3_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MI3_OK;
}

int mi3_batch_destroy(mi3_batch* b) {
    if (!b) return MI3_EINVAL;
    hipSetDevice(b->dev);
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto& kv : b->classes) {
        if (kv.second.side) {
            hipStreamSynchronize(kv.second.side);
            hipStreamDestroy(kv.second.side);
        }
        hipFree(kv.second.tiles);
    }
    for (uint32_t m = 0; m < b->n; m++) {
        hipFree(b->Jp[m]);
        hipFree(b->Ju[m]);
        hipFree(b->seq[m]);
        hipFree(b->pos[m]);
    }
    hipFree(b->models);
    hipFree(b->pos_ptrs);
    hipFree(b->steps);
    hipFree(b->counters);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return MI3_OK;
}

const char* mi3_batch_last_error(mi3_batch* b) {
    return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int mi3_batch_set_couplings(mi3_batch* b, uint32_t model, const float* J) {
    if (!b || !J) return MI3_EINVAL;
    if (model >= b->n) return bfail(b, MI3_EINVAL, "model index out of range");
    hipSetDevice(b->dev);
    const size_t np = (size_t)b->L[model] * (b->L[model] - 1) / 2 * b->q[model] * b->q[model];
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(b->Jp[model], J, np * 4, hipMemcpyHostToDevice));
    BCHK(b, launch_unpack_chunked(b->Jp[model], b->Ju[model], b->L[model], b->q[model],
                                  b->nchunk[model], b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return MI3_OK;
}

int mi3_batch_set_seqs(mi3_batch* b, uint32_t model, const uint8_t* seqs) {
    if (!b || !seqs) return MI3_EINVAL;
    if (model >= b->n) return bfail(b, MI3_EINVAL, "model index out of range");
    hipSetDevice(b->dev);
    const size_t nb = (size_t)b->L[model] * b->N[model];
    for (size_t i = 0; i < nb; i++)
        if (seqs[i] >= b->q[model]) return bfail(b, MI3_EINVAL, "residues must be < q");
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(b->seq[model], seqs, nb, hipMemcpyHostToDevice));
    return MI3_OK;
}

int mi3_batch_get_seqs(mi3_batch* b, uint32_t model, uint8_t* out) {
    if (!b || !out) return MI3_EINVAL;
    if (model >= b->n) return bfail(b, MI3_EINVAL, "model index out of range");
    hipSetDevice(b->dev);
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(out, b->seq[model], (size_t)b->L[model] * b->N[model],
                      hipMemcpyDeviceToHost));
    return MI3_OK;
}

int mi3_batch_run(mi3_batch* b, uint32_t nsweeps) {
    if (!b || nsweeps == 0) return MI3_EINVAL;
    hipSetDevice(b->dev);
    if (!b->exec || b->nsweeps != nsweeps) {
        drop_graph(b);
        BCHK(b, hipStreamSynchronize(b->stream));
        int rc = upload_models(b, nsweeps);
        if (rc) return rc;
        BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        std::vector<hipEvent_t> events;   // fork/join events, alive until capture ends
        rc = enqueue_round(b, events);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        for (auto ev : events) hipEventDestroy(ev);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return bfail(b, MI3_EHIP, std::string("capture: ") + hipGetErrorString(e));
        b->graph = g;
        BCHK(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->nsweeps = nsweeps;
    }
    BCHK(b, hipGraphLaunch(b->exec, b->stream));
    return MI3_OK;
}

int mi3_batch_synchronize(mi3_batch* b) {
    if (!b) return MI3_EINVAL;
    hipSetDevice(b->dev);
    BCHK(b, hipStreamSynchronize(b->stream));
    return MI3_OK;
}

int mi3_batch_tune(mi3_batch* b, uint32_t nsweeps) {
    if (!b || nsweeps == 0) return MI3_EINVAL;
    hipSetDevice(b->dev);
    drop_graph(b);
    BCHK(b, hipStreamSynchronize(b->stream));
    int rc = upload_models(b, nsweeps);
    if (rc) return rc;
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    uint32_t ci = 0;
    for (auto& kv : b->classes) {
        ClassRun& cr = kv.second;
        float best = 1e30f;
        uint32_t best_w = cr.wpb;
        for (uint32_t w : tile_options(kv.first.kw)) {
            rc = build_tiles(b, cr, w);
            if (rc) break;
            // positions for this launch, then the class sweep alone; the
            // step counters advance as after any launch
            hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, b->stream, b->counters + ci, 1u);
            const uint32_t mx = max_steps(b);
            hipLaunchKernelGGL(k_batch_positions,
                               dim3(std::max(1u, std::min(64u, (mx + 255) / 256)), b->n), dim3(256),
                               0, b->stream, (const BatchModel*)b->models,
                               (uint32_t* const*)b->pos_ptrs);
            hipEventRecord(e0, b->stream);
            hipError_t e = launch_sweep_batch(kv.first.kw, kv.first.pf, w, b->models, cr.tiles,
                                              cr.ntiles, b->counters + ci, class_grid(b, cr),
                                              b->stream);
            hipEventRecord(e1, b->stream);
            hipLaunchKernelGGL(k_batch_advance, dim3((b->n + 255) / 256), dim3(256), 0, b->stream,
                               (const BatchModel*)b->models, b->steps, b->n);
            if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = bfail(b, MI3_EHIP, "tuning launch failed");
                break;
            }
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                best_w = w;
            }
        }
        if (rc) break;
        rc = build_tiles(b, cr, best_w);
        if (rc) break;
        ci++;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int mi3_batch_tile(mi3_batch* b, uint32_t model, uint32_t* wpb) {
    if (!b || !wpb || model >= b->n) return MI3_EINVAL;
    ClassKey k{state_words(b->L[model]), pf_class(b->q[model])};
    *wpb = b->classes[k].wpb;
    return MI3_OK;
}

int mi3_batch_set_tile(mi3_batch* b, uint32_t model, uint32_t wpb) {
    if (!b || model >= b->n) return MI3_EINVAL;
    ClassKey k{state_words(b->L[model]), pf_class(b->q[model])};
    auto opts = tile_options(k.kw);
    if (std::find(opts.begin(), opts.end(), wpb) == opts.end())
        return bfail(b, MI3_EINVAL, "walkers per tile must be 128, 256 or 512 (128 or 256 for L > 512)");
    hipSetDevice(b->dev);
    BCHK(b, hipStreamSynchronize(b->stream));
    drop_graph(b);
    return build_tiles(b, b->classes[k], wpb);
}

int mi3_batch_step_count(mi3_batch* b, uint32_t model, uint64_t* steps) {
    if (!b || !steps || model >= b->n) return MI3_EINVAL;
    hipSetDevice(b->dev);
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(steps, b->steps + model, sizeof(uint64_t), hipMemcpyDeviceToHost));
    return MI3_OK;
}

}  // extern "C"
