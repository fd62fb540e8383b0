// libmi3hip.so: context, named buffers, stream, tickets, profiling and the
// C ABI declared in include/mi3hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mi3hip.h"
#include "launch.h"
#include "k_misc.hip"
#include "k_newton.hip"
#include "k_bicount.hip"

using namespace mi3;

// Library-owned scratch of the allocation-free entry point (one copy per
// device; see mi3_weighted_marg_noalloc).
static __device__ __align__(256) unsigned char scratch[MI3_NOALLOC_SCRATCH_BYTES];

namespace mi3 {
hipError_t launch_unpack_chunked(const float* J, float* out, uint32_t L, uint32_t q,
                                 uint32_t nchunk, hipStream_t st) {
    const uint32_t chf = chunk_floats(q);
    const uint64_t total = (uint64_t)L * nchunk * chf;
    hipLaunchKernelGGL(k_unpack_chunked, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st,
                       J, out, L, q, nchunk, chf, total);
    return hipGetLastError();
}
}  // namespace mi3

namespace {

struct Buf {
    void* d = nullptr;
    size_t bytes = 0;
    uint64_t version = 0;
};

struct Pinned {
    void* h = nullptr;
    size_t cap = 0;
    bool busy = false;
};

struct Ticket {
    int pin = -1;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool live = false;
};

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
    size_t bytes;
};

struct Chunked {
    float* d = nullptr;
    uint64_t version = ~0ull;
};

thread_local std::string g_create_error;

}  // namespace

struct mi3_ctx {
    int dev = 0;
    uint32_t L = 0, q = 0, nP = 0, N = 0, nchunk = 0;
    uint64_t seed = 0;
    uint32_t k0 = 0, k1 = 0;
    hipStream_t stream = nullptr;
    std::map<std::string, Buf> bufs;
    std::string err;
    // MCMC
    bool mcmc = false;
    uint32_t nsteps = 0;
    uint64_t step = 0;
    uint32_t walker_base = 0;
    uint32_t indep_calls = 0;
    uint32_t* pos = nullptr;
    size_t pos_cap = 0;
    uint32_t wpb = 0;  // sweep walkers per work-group (0 = default)
    uint32_t ncu = 256;  // compute units of the device
    // derived device data
    std::map<std::string, Chunked> chunked;  // chunked layouts keyed by source buffer
    double* red = nullptr;                    // reduction partials
    float* red_out = nullptr;                 // 2 floats: internal reduction result
    uint64_t* wq = nullptr;                   // fixed-point weights
    size_t wq_cap = 0;
    unsigned long long* wacc = nullptr;       // fixed-point weighted-marginal bins
    unsigned long long* probe_out = nullptr;  // sink of mi3_wm_atomic_probe
    uint32_t* wbad = nullptr;
    uint32_t* wm_units = nullptr;
    uint32_t n_units = 0;
    uint32_t wm_split = 1;
    std::vector<uint32_t> unit_pair_lo, unit_pair_hi;  // pair range [lo, hi) of each unit
    // this device's copy of the allocation-free scratch array (null, with
    // the reason in scratch_err, if its address could not be resolved)
    char* scratch = nullptr;
    std::string scratch_err;
    uint8_t* seq1 = nullptr;                  // one sequence (fill_seqs)
    // host staging
    std::vector<Pinned> pins;
    std::vector<Ticket> tickets;
    // profiling
    bool profile = false;
    hipEvent_t t0 = nullptr;  // time origin of EVT records
    std::vector<ProfRec> prof;
};

namespace {

constexpr uint32_t RED_BLOCKS = 512;

// Scratch of mi3_weighted_marg_noalloc: one static device array per GPU the
// library runs on, held until the process ends.  Layout: fixed-point weights
// (8 B per sequence), reduction partials, the reduced (sum, sum of squares),
// the bad-total flag, and a uint64 accumulator for a run of pair blocks.
constexpr size_t NA_WQ_BYTES = (size_t)MI3_NOALLOC_MAX_SEQ * 8 + 256;
constexpr size_t NA_RED_OFF = NA_WQ_BYTES;
constexpr size_t NA_REDOUT_OFF = NA_RED_OFF + RED_BLOCKS * 2 * sizeof(double);
constexpr size_t NA_BAD_OFF = NA_REDOUT_OFF + 256;
constexpr size_t NA_ACC_OFF = NA_BAD_OFF + 256;
constexpr uint64_t NA_ACC_BINS = 1ull << 21;
constexpr size_t NA_BYTES = NA_ACC_OFF + NA_ACC_BINS * 8;
static_assert(NA_BYTES == MI3_NOALLOC_SCRATCH_BYTES, "scratch size declared in mi3hip.h");

// Device address of `scratch` on each device, resolved once per device with
// that device current (the array's name in host code is only a stand-in).
constexpr int NA_MAX_DEV = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[NA_MAX_DEV] = {};

int fail(mi3_ctx* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_create_error = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(ctx, MI3_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

Buf* find(mi3_ctx* c, const char* name) {
    if (!name) return nullptr;
    auto it = c->bufs.find(name);
    return it == c->bufs.end() ? nullptr : &it->second;
}

int need(mi3_ctx* c, const char* name, size_t bytes, Buf** out) {
    Buf* b = find(c, name);
    if (!b) return fail(c, MI3_ENOBUF, std::string("no buffer '") + (name ? name : "(null)") + "'");
    if (b->bytes < bytes)
        return fail(c, MI3_EINVAL,
                    std::string("buffer '") + name + "' holds " + std::to_string(b->bytes) +
                        " bytes, " + std::to_string(bytes) + " needed");
    *out = b;
    return MI3_OK;
}

// Row stride (bytes) of a site-major seq buffer.
size_t seq_ld(mi3_ctx* c, Buf* b) { return c->L ? b->bytes / c->L : 0; }

struct Scope {
    mi3_ctx* c;
    const char* name;
    size_t bytes;
    hipEvent_t a = nullptr;
    Scope(mi3_ctx* c_, const char* n, size_t by) : c(c_), name(n), bytes(by) {
        if (c->profile) {
            hipEventCreate(&a);
            hipEventRecord(a, c->stream);
        }
    }
    ~Scope() {
        if (c->profile) {
            hipEvent_t b;
            hipEventCreate(&b);
            hipEventRecord(b, c->stream);
            c->prof.push_back({name, a, b, bytes});
        }
    }
};

uint32_t grid1(uint64_t n, uint32_t bs = 256) { return (uint32_t)((n + bs - 1) / bs); }

int fill_u32_dev(mi3_ctx* c, void* p, uint32_t val, uint64_t nwords) {
    if (nwords == 0) return MI3_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid1(nwords)), dim3(256), 0, c->stream, (uint32_t*)p, val,
                       nwords);
    HIPCHK(c, hipGetLastError());
    return MI3_OK;
}

// Chunked symmetric copy of a packed coupling buffer, rebuilt when stale.
int chunked_of(mi3_ctx* c, const char* name, float** out) {
    Buf* b;
    size_t qq = (size_t)c->q * c->q;
    int rc = need(c, name, (size_t)c->nP * qq * sizeof(float), &b);
    if (rc) return rc;
    Chunked& ch = c->chunked[name];
    const uint32_t chf = chunk_floats(c->q);
    uint64_t total = (uint64_t)c->L * c->nchunk * chf;
    if (!ch.d) HIPCHK(c, hipMalloc(&ch.d, total * sizeof(float)));
    if (ch.version != b->version) {
        Scope s(c, "unpack_chunked", total * 4);
        hipLaunchKernelGGL(k_unpack_chunked, dim3(grid1(total)), dim3(256), 0, c->stream,
                           (const float*)b->d, ch.d, c->L, c->q, c->nchunk, chf, total);
        HIPCHK(c, hipGetLastError());
        ch.version = b->version;
    }
    *out = ch.d;
    return MI3_OK;
}

int pin_get(mi3_ctx* c, size_t bytes) {
    for (size_t i = 0; i < c->pins.size(); i++)
        if (!c->pins[i].busy && c->pins[i].cap >= bytes) {
            c->pins[i].busy = true;
            return (int)i;
        }
    Pinned p;
    p.cap = bytes < 4096 ? 4096 : bytes;
    if (hipHostMalloc(&p.h, p.cap, hipHostMallocDefault) != hipSuccess) return -1;
    p.busy = true;
    c->pins.push_back(p);
    return (int)c->pins.size() - 1;
}

// Parts of the sequence range for the bicount kernel: about four waves per
// SIMD (one wave fits per SIMD at a time) share the work evenly, keeping at
// least 64 k-steps of 32 sequences per wave; each wave's int32 accumulators
// reach -128 per co-occurrence, so a part stays below 2^24 sequences
// (|acc| < 2^31).
uint32_t bicount_split(const mi3_ctx* c, uint32_t nseq) {
    const uint32_t Lq = c->L * c->q;
    const uint32_t nblk = (Lq + 127) / 128;
    const uint32_t nbw = nblk * (nblk + 1) / 2;
    const uint32_t ksteps = (nseq + 31) / 32;
    const uint32_t target = 4 * 4 * c->ncu;
    uint32_t ksplit = (target + nbw / 2) / nbw;
    const uint32_t kmax = ksteps / 64 > 1 ? ksteps / 64 : 1;
    if (ksplit > kmax) ksplit = kmax;
    if (ksplit > 32) ksplit = 32;
    if (ksplit < 1) ksplit = 1;
    while ((uint64_t)((ksteps + ksplit - 1) / ksplit) * 32 >= (1ull << 24)) ksplit++;
    return ksplit;
}

}  // namespace

extern "C" {

const char* mi3_version(void) { return "mi3hip 0.1 (gfx950)"; }

int mi3_device_count(int* count) {
    if (!count) return MI3_EINVAL;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return fail(nullptr, MI3_EHIP, hipGetErrorString(e));
    }
    return MI3_OK;
}

int mi3_device_info(int device, char* buf, size_t buflen) {
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, device);
    if (e != hipSuccess) return fail(nullptr, MI3_EHIP, hipGetErrorString(e));
    snprintf(buf, buflen,
             "  Device '%s' (%s):\n    Compute Units: %d\n    Clock (kHz): %d\n"
             "    Global Mem Size: %zu\n    L2 Cache Size: %d\n    LDS per block: %zu\n",
             p.name, p.gcnArchName, p.multiProcessorCount, p.clockRate, (size_t)p.totalGlobalMem,
             p.l2CacheSize, (size_t)p.sharedMemPerBlock);
    return MI3_OK;
}

int mi3_create(int device, uint32_t L, uint32_t q, uint32_t nwalkers, uint64_t seed, int profile,
               mi3_ctx** out) {
    if (!out) return fail(nullptr, MI3_EINVAL, "out is NULL");
    *out = nullptr;
    if (L < 2 || q < 2) return fail(nullptr, MI3_EINVAL, "need L >= 2 and q >= 2");
    if (q > MAX_Q || L > MAX_L)
        return fail(nullptr, MI3_EUNSUP,
                    "this build supports q <= " + std::to_string(MAX_Q) + " and L <= " +
                        std::to_string(MAX_L));
    if (!chunked_fits(L, q))
        return fail(nullptr, MI3_EUNSUP,
                    "L = " + std::to_string(L) + " with q = " + std::to_string(q) +
                        ": the chunked couplings would exceed 4 GiB");
    if (nwalkers % 512 != 0)
        return fail(nullptr, MI3_EINVAL, "nwalkers/ngpus must be a multiple of 512");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, MI3_EHIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(nullptr, MI3_EINVAL, "no such device");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, MI3_EHIP, "hipSetDevice failed");
    mi3_ctx* c = new mi3_ctx;
    c->dev = device;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->ncu = (uint32_t)prop.multiProcessorCount;
    }
    c->L = L;
    c->q = q;
    c->nP = L * (L - 1) / 2;
    c->N = nwalkers;
    c->nchunk = (L + CH - 1) / CH;
    c->seed = seed;
    c->k0 = (uint32_t)seed;
    c->k1 = (uint32_t)(seed >> 32);
    c->profile = profile != 0;
    {
        // this device's address of the scratch array, resolved once per
        // device with the device current
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= NA_MAX_DEV) {
            c->scratch_err = "allocation-free scratch: device index beyond the table";
        } else if (!g_scratch_ptr[device]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e == hipSuccess && p)
                g_scratch_ptr[device] = static_cast<char*>(p);
            else
                c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") +
                                 hipGetErrorString(e);
        }
        if (device < NA_MAX_DEV) c->scratch = g_scratch_ptr[device];
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, MI3_EHIP, "stream creation failed");
    }
    if (c->profile) {
        hipEventCreate(&c->t0);
        hipEventRecord(c->t0, c->stream);
    }
    // weighted-marginal work units: 32 consecutive pairs of the packed order
    // each (the last one may be short); each unit's sequence range is split
    // in wm_split parts, chosen to fill the last wave of work-groups best
    std::vector<uint32_t> units;
    const uint32_t npairs = L * (L - 1) / 2;
    for (uint32_t u = 0; u * WM_COLS < npairs; u++) {
        units.push_back(u);
        c->unit_pair_lo.push_back(u * WM_COLS);
        c->unit_pair_hi.push_back(std::min(npairs, (u + 1) * WM_COLS));
    }
    c->n_units = (uint32_t)units.size();
    {
        double best = 0.0;
        for (uint32_t ns = 1; ns <= 4; ns++) {
            const uint32_t wgs = c->n_units * ns;
            const double t = (double)((wgs + c->ncu - 1) / c->ncu) / ns;
            if (ns == 1 || t < best - 1e-9) {
                best = t;
                c->wm_split = ns;
            }
        }
    }
    if (hipMalloc(&c->wm_units, units.size() * 4) != hipSuccess ||
        hipMemcpy(c->wm_units, units.data(), units.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c->red, RED_BLOCKS * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->red_out, 2 * sizeof(float)) != hipSuccess ||
        hipMalloc(&c->wbad, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&c->seq1, L) != hipSuccess) {
        mi3_destroy(c);
        return fail(nullptr, MI3_ENOMEM, "device allocation failed");
    }
    *out = c;
    return MI3_OK;
}

int mi3_destroy(mi3_ctx* c) {
    if (!c) return MI3_OK;
    hipSetDevice(c->dev);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& kv : c->bufs) hipFree(kv.second.d);
    for (auto& kv : c->chunked) hipFree(kv.second.d);
    for (auto& p : c->pins) hipHostFree(p.h);
    for (auto& t : c->tickets)
        if (t.ev) hipEventDestroy(t.ev);
    for (auto& r : c->prof) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    hipFree(c->pos);
    hipFree(c->red);
    hipFree(c->red_out);
    hipFree(c->wq);
    hipFree(c->wacc);
    hipFree(c->probe_out);
    hipFree(c->wbad);
    hipFree(c->wm_units);
    hipFree(c->seq1);
    if (c->t0) hipEventDestroy(c->t0);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return MI3_OK;
}

const char* mi3_last_error(mi3_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int mi3_buf_alloc(mi3_ctx* c, const char* name, size_t nbytes) {
    if (!c || !name) return MI3_EINVAL;
    if (find(c, name)) return fail(c, MI3_ESTATE, std::string("buffer '") + name + "' exists");
    Buf b;
    b.bytes = nbytes;
    // pad so that 16-byte vector loads at the tail stay in bounds
    size_t alloc = ((nbytes + 255) / 256) * 256 + 256;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMalloc(&b.d, alloc));
    int rc = fill_u32_dev(c, b.d, 0u, alloc / 4);
    if (rc) return rc;
    c->bufs[name] = b;
    return MI3_OK;
}

int mi3_buf_exists(mi3_ctx* c, const char* name) { return (c && find(c, name)) ? 1 : 0; }

int mi3_buf_device_ptr(mi3_ctx* c, const char* name, void** ptr) {
    if (!c || !ptr) return MI3_EINVAL;
    Buf* b = find(c, name);
    if (!b) return fail(c, MI3_ENOBUF, std::string("no buffer '") + (name ? name : "(null)") + "'");
    *ptr = b->d;
    return MI3_OK;
}

int mi3_buf_mark_written(mi3_ctx* c, const char* name) {
    if (!c) return MI3_EINVAL;
    Buf* b = find(c, name);
    if (!b) return fail(c, MI3_ENOBUF, std::string("no buffer '") + (name ? name : "(null)") + "'");
    b->version++;
    return MI3_OK;
}

int mi3_stream(mi3_ctx* c, void** stream) {
    if (!c || !stream) return MI3_EINVAL;
    *stream = c->stream;
    return MI3_OK;
}

int mi3_buf_size(mi3_ctx* c, const char* name, size_t* nbytes) {
    Buf* b = c ? find(c, name) : nullptr;
    if (!b) return c ? fail(c, MI3_ENOBUF, "no such buffer") : MI3_EINVAL;
    *nbytes = b->bytes;
    return MI3_OK;
}

int mi3_buf_write(mi3_ctx* c, const char* name, const void* src, size_t nbytes, size_t offset) {
    if (!c || (!src && nbytes)) return MI3_EINVAL;
    Buf* b;
    int rc = need(c, name, offset + nbytes, &b);
    if (rc) return rc;
    if (nbytes == 0) return MI3_OK;
    Scope s(c, "setBuf", nbytes);
    HIPCHK(c, hipMemcpyAsync((char*)b->d + offset, src, nbytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b->version++;
    return MI3_OK;
}

int mi3_buf_copy(mi3_ctx* c, const char* dst, const char* src, size_t nbytes) {
    if (!c) return MI3_EINVAL;
    Buf *bd, *bs;
    int rc = need(c, dst, nbytes, &bd);
    if (rc) return rc;
    rc = need(c, src, nbytes, &bs);
    if (rc) return rc;
    Scope s(c, "setBuf", nbytes);
    HIPCHK(c, hipMemcpyAsync(bd->d, bs->d, nbytes, hipMemcpyDeviceToDevice, c->stream));
    bd->version++;
    return MI3_OK;
}

int mi3_stream_copy(mi3_ctx* c, const char* dst, const char* src, size_t nbytes) {
    if (!c) return MI3_EINVAL;
    if (nbytes % 16) return fail(c, MI3_EINVAL, "stream copy: nbytes must be a multiple of 16");
    Buf *bd, *bs;
    int rc = need(c, dst, nbytes, &bd);
    if (rc) return rc;
    rc = need(c, src, nbytes, &bs);
    if (rc) return rc;
    Scope s(c, "streamCopy", 2 * nbytes);
    hipLaunchKernelGGL(k_stream_copy, dim3(c->ncu * 256), dim3(256), 0, c->stream,
                       (const u4v*)bs->d, (u4v*)bd->d, (uint64_t)(nbytes / 16));
    HIPCHK(c, hipGetLastError());
    bd->version++;
    return MI3_OK;
}

int mi3_stream_read(mi3_ctx* c, const char* src, size_t nbytes, const char* out) {
    if (!c) return MI3_EINVAL;
    if (nbytes % 16) return fail(c, MI3_EINVAL, "stream read: nbytes must be a multiple of 16");
    Buf *bs, *bo;
    int rc = need(c, src, nbytes, &bs);
    if (rc) return rc;
    rc = need(c, out, 4, &bo);
    if (rc) return rc;
    Scope s(c, "streamRead", nbytes);
    hipLaunchKernelGGL(k_stream_read, dim3(c->ncu * 256), dim3(256), 0, c->stream,
                       (const u4v*)bs->d, (uint64_t)(nbytes / 16), 0x5A5A5A5Au,
                       (uint32_t*)bo->d);
    HIPCHK(c, hipGetLastError());
    return MI3_OK;
}

int mi3_buf_read_async(mi3_ctx* c, const char* name, size_t nbytes, size_t offset, int* ticket) {
    if (!c || !ticket) return MI3_EINVAL;
    Buf* b;
    int rc = need(c, name, offset + nbytes, &b);
    if (rc) return rc;
    int pin = pin_get(c, nbytes);
    if (pin < 0) return fail(c, MI3_ENOMEM, "pinned host allocation failed");
    Ticket t;
    t.pin = pin;
    t.bytes = nbytes;
    t.live = true;
    HIPCHK(c, hipEventCreateWithFlags(&t.ev, hipEventDisableTiming));
    {
        Scope s(c, "getBuf", nbytes);
        if (nbytes)
            HIPCHK(c, hipMemcpyAsync(c->pins[pin].h, (char*)b->d + offset, nbytes,
                                     hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipEventRecord(t.ev, c->stream));
    for (size_t i = 0; i < c->tickets.size(); i++)
        if (!c->tickets[i].live) {
            c->tickets[i] = t;
            *ticket = (int)i;
            return MI3_OK;
        }
    c->tickets.push_back(t);
    *ticket = (int)c->tickets.size() - 1;
    return MI3_OK;
}

int mi3_ticket_wait(mi3_ctx* c, int ticket, void* dst, size_t nbytes) {
    if (!c || ticket < 0 || (size_t)ticket >= c->tickets.size() || !c->tickets[ticket].live)
        return c ? fail(c, MI3_EINVAL, "invalid ticket") : MI3_EINVAL;
    Ticket& t = c->tickets[ticket];
    HIPCHK(c, hipEventSynchronize(t.ev));
    size_t n = nbytes < t.bytes ? nbytes : t.bytes;
    if (dst && n) memcpy(dst, c->pins[t.pin].h, n);
    c->pins[t.pin].busy = false;
    hipEventDestroy(t.ev);
    t.ev = nullptr;
    t.live = false;
    return MI3_OK;
}

int mi3_buf_fill_u32(mi3_ctx* c, const char* name, uint32_t pattern) {
    if (!c) return MI3_EINVAL;
    Buf* b;
    int rc = need(c, name, 0, &b);
    if (rc) return rc;
    Scope s(c, "fill_buffer", b->bytes);
    rc = fill_u32_dev(c, b->d, pattern, (b->bytes + 3) / 4);
    if (rc) return rc;
    b->version++;
    return MI3_OK;
}

int mi3_synchronize(mi3_ctx* c) {
    if (!c) return MI3_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MI3_OK;
}

int mi3_fill_seqs(mi3_ctx* c, const char* seqbuf, const uint8_t* seq) {
    if (!c || !seq) return MI3_EINVAL;
    Buf* b;
    int rc = need(c, seqbuf, (size_t)c->L, &b);
    if (rc) return rc;
    size_t ld = seq_ld(c, b);
    HIPCHK(c, hipMemcpyAsync(c->seq1, seq, c->L, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Scope s(c, "fillSeqs", b->bytes);
    uint64_t n = (uint64_t)c->L * ld;
    hipLaunchKernelGGL(k_fill_seqs, dim3(grid1(n)), dim3(256), 0, c->stream, (uint8_t*)b->d,
                       (uint32_t)ld, (uint32_t)ld, (const uint8_t*)c->seq1, c->L);
    HIPCHK(c, hipGetLastError());
    b->version++;
    return MI3_OK;
}

int mi3_copy_seqs(mi3_ctx* c, const char* src, uint32_t src_offset, const char* dst,
                  uint32_t dst_offset, uint32_t n) {
    if (!c) return MI3_EINVAL;
    Buf *bs, *bd;
    int rc = need(c, src, (size_t)c->L, &bs);
    if (rc) return rc;
    rc = need(c, dst, (size_t)c->L, &bd);
    if (rc) return rc;
    const size_t lds = seq_ld(c, bs), ldd = seq_ld(c, bd);
    if ((size_t)src_offset + n > lds || (size_t)dst_offset + n > ldd)
        return fail(c, MI3_EINVAL, "sequence range past the end of a seq buffer");
    if (n == 0) return MI3_OK;
    Scope s(c, "copySeqs", (size_t)2 * n * c->L);
    HIPCHK(c, hipMemcpy2DAsync((uint8_t*)bd->d + dst_offset, ldd, (const uint8_t*)bs->d + src_offset,
                               lds, n, c->L, hipMemcpyDeviceToDevice, c->stream));
    bd->version++;
    return MI3_OK;
}

int mi3_copy_subseq(mi3_ctx* c, const char* src, uint32_t seqind, const char* dst,
                    const char* marks) {
    if (!c) return MI3_EINVAL;
    Buf *bs, *bd, *bm;
    int rc = need(c, src, (size_t)c->L, &bs);
    if (rc) return rc;
    rc = need(c, dst, (size_t)c->L, &bd);
    if (rc) return rc;
    rc = need(c, marks, (size_t)c->L, &bm);
    if (rc) return rc;
    const size_t lds = seq_ld(c, bs), ldd = seq_ld(c, bd);
    if (seqind >= lds) return fail(c, MI3_EINVAL, "given index is past end of the seq buffer");
    uint64_t n = (uint64_t)c->L * ldd;
    Scope s(c, "copySubseq", n);
    hipLaunchKernelGGL(k_copy_subseq, dim3(grid1(n)), dim3(256), 0, c->stream, (uint8_t*)bd->d,
                       (uint32_t)ldd, (uint32_t)ldd, (const uint8_t*)bs->d, (uint32_t)lds, seqind,
                       (const uint8_t*)bm->d, c->L);
    HIPCHK(c, hipGetLastError());
    bd->version++;
    return MI3_OK;
}

int mi3_set_wgsize(mi3_ctx* c, uint32_t wgsize) {
    if (!c) return MI3_EINVAL;
    if (wgsize == 0 || (wgsize & (wgsize - 1)))
        return fail(c, MI3_EINVAL, "wgsize must be a power of two");
    c->wpb = wgsize <= 128 ? 128 : (wgsize <= 256 ? 256 : 512);
    return MI3_OK;
}

int mi3_init_mcmc(mi3_ctx* c, uint32_t nsteps, uint64_t rng_offset, uint64_t rng_span) {
    if (!c) return MI3_EINVAL;
    if (c->mcmc) return fail(c, MI3_ESTATE, "Already initialized MCMC");
    if (nsteps == 0) return fail(c, MI3_EINVAL, "nsteps must be positive");
    if (c->pos_cap < nsteps) {
        hipFree(c->pos);
        c->pos = nullptr;
        c->pos_cap = 0;
        HIPCHK(c, hipMalloc(&c->pos, (size_t)nsteps * 4));
        c->pos_cap = nsteps;
    }
    c->nsteps = nsteps;
    c->walker_base = rng_span ? (uint32_t)(rng_offset / rng_span) * c->N : 0u;
    c->mcmc = true;
    return MI3_OK;
}

int mi3_run_mcmc(mi3_ctx* c, float beta) {
    if (!c) return MI3_EINVAL;
    if (!c->mcmc) return fail(c, MI3_ESTATE, "MCMC not initialized");
    Buf* sb;
    int rc = need(c, "seq main", (size_t)c->L * c->N, &sb);
    if (rc) return rc;
    float* Ju;
    rc = chunked_of(c, "J", &Ju);
    if (rc) return rc;
    {
        Scope s(c, "positions", (size_t)c->nsteps * 4);
        hipLaunchKernelGGL(k_positions, dim3(grid1(c->nsteps)), dim3(256), 0, c->stream, c->pos,
                           c->nsteps, c->L, c->k0, c->k1, c->step);
        HIPCHK(c, hipGetLastError());
    }
    SweepArgs a;
    a.Ju = Ju;
    a.seq = (uint8_t*)sb->d;
    a.pos = c->pos;
    a.ld = (uint32_t)seq_ld(c, sb);
    a.nwalk = c->N;
    a.L = c->L;
    a.q = c->q;
    a.nchunk = c->nchunk;
    a.nsteps = c->nsteps;
    a.k0 = c->k0;
    a.k1 = c->k1;
    a.walker_base = c->walker_base;
    a.step0 = c->step;
    a.nbeta = -beta;
    a.wpb = c->wpb;
    {
        Scope s(c, "mcmc", (size_t)c->N * c->nsteps * (c->L + 1));
        HIPCHK(c, launch_sweep(a, c->stream));
    }
    c->step += c->nsteps + (c->nsteps & 1u);  // keep the global step even
    sb->version++;
    return MI3_OK;
}

int mi3_wm_atomic_probe(mi3_ctx* c, uint32_t work_groups, uint32_t iters) {
    if (!c) return MI3_EINVAL;
    if (work_groups == 0 || work_groups > 65536 || iters == 0 || iters > (1u << 20))
        return fail(c, MI3_EINVAL, "work_groups must be 1..65536 and iters 1..2^20");
    if (!c->probe_out) HIPCHK(c, hipMalloc(&c->probe_out, sizeof(unsigned long long)));
    const size_t lds = (size_t)c->q * c->q * WM_COLS * sizeof(unsigned long long);
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_wm_atomic_probe),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    Scope s(c, "wmAtomicProbe", (size_t)work_groups * WM_THREADS * iters * 8);
    hipLaunchKernelGGL(k_wm_atomic_probe, dim3(work_groups), dim3(WM_THREADS), lds, c->stream,
                       c->q, iters, c->probe_out);
    HIPCHK(c, hipGetLastError());
    return MI3_OK;
}

int mi3_stage_probe(mi3_ctx* c, int inflight) {
    if (!c) return MI3_EINVAL;
    if (!c->mcmc) return fail(c, MI3_ESTATE, "MCMC not initialized");
    if (inflight != 1 && inflight != 2) return fail(c, MI3_EINVAL, "inflight must be 1 or 2");
    float* Ju;
    int rc = chunked_of(c, "J", &Ju);
    if (rc) return rc;
    {
        Scope s(c, "positions", (size_t)c->nsteps * 4);
        hipLaunchKernelGGL(k_positions, dim3(grid1(c->nsteps)), dim3(256), 0, c->stream, c->pos,
                           c->nsteps, c->L, c->k0, c->k1, c->step);
        HIPCHK(c, hipGetLastError());
    }
    SweepArgs a{};
    a.Ju = Ju;
    a.pos = c->pos;
    a.nwalk = c->N;
    a.L = c->L;
    a.q = c->q;
    a.nchunk = c->nchunk;
    a.nsteps = c->nsteps;
    const size_t bytes = (size_t)(c->N / 512) * c->nsteps * c->nchunk * chunk_floats(c->q) * 4;
    Scope s(c, inflight == 1 ? "stageProbe" : "stageProbe2", bytes);
    HIPCHK(c, launch_stage_probe(a, inflight, c->stream));
    return MI3_OK;
}

int mi3_mcmc_step_count(mi3_ctx* c, uint64_t* steps) {
    if (!c || !steps) return MI3_EINVAL;
    *steps = c->step;
    return MI3_OK;
}

int mi3_gen_indep(mi3_ctx* c, const char* seqbuf) {
    if (!c) return MI3_EINVAL;
    Buf *sb, *cp;
    int rc = need(c, seqbuf, (size_t)c->L, &sb);
    if (rc) return rc;
    rc = need(c, "cprob", (size_t)c->L * (c->q - 1) * 4, &cp);
    if (rc) return rc;
    uint32_t ld = (uint32_t)seq_ld(c, sb);
    Scope s(c, "gen_indep", sb->bytes);
    hipLaunchKernelGGL(k_gen_indep, dim3(grid1(ld)), dim3(256), 0, c->stream, (uint8_t*)sb->d, ld,
                       ld, (const float*)cp->d, c->L, c->q, c->k0, c->k1, c->indep_calls,
                       c->walker_base);
    HIPCHK(c, hipGetLastError());
    c->indep_calls++;
    sb->version++;
    return MI3_OK;
}

int mi3_calc_energies(mi3_ctx* c, const char* seqbuf, const char* Jbuf, const char* Ebuf,
                      uint32_t nseq) {
    if (!c) return MI3_EINVAL;
    Buf *sb, *eb;
    int rc = need(c, seqbuf, (size_t)c->L, &sb);
    if (rc) return rc;
    size_t ld = seq_ld(c, sb);
    if (nseq > ld) return fail(c, MI3_EINVAL, "nseq exceeds the sequence buffer");
    rc = need(c, Ebuf, (size_t)nseq * 4, &eb);
    if (rc) return rc;
    float* Ju;
    rc = chunked_of(c, Jbuf, &Ju);
    if (rc) return rc;
    EnergyArgs a;
    a.Ju = Ju;
    a.seq = (const uint8_t*)sb->d;
    a.E = (float*)eb->d;
    a.ld = (uint32_t)ld;
    a.nseq = nseq;
    a.L = c->L;
    a.q = c->q;
    a.nchunk = c->nchunk;
    Scope s(c, "getEnergies", (size_t)nseq * (c->L + 4));
    HIPCHK(c, launch_energies(a, c->stream));
    eb->version++;
    return MI3_OK;
}

int mi3_calc_bicounts(mi3_ctx* c, const char* seqbuf, uint32_t nseq) {
    if (!c) return MI3_EINVAL;
    Buf *sb, *cb;
    int rc = need(c, seqbuf, (size_t)c->L, &sb);
    if (rc) return rc;
    size_t ld = seq_ld(c, sb);
    if (nseq > ld) return fail(c, MI3_EINVAL, "nseq exceeds the sequence buffer");
    if (ld % 16) return fail(c, MI3_EINVAL, "sequence buffer row stride must be a multiple of 16");
    if ((uint64_t)ld * c->L >= (1ull << 31))
        return fail(c, MI3_EUNSUP, "bicounts: sequence buffer of 2 GiB or more");
    rc = need(c, "bicount", (size_t)c->nP * c->q * c->q * 4, &cb);
    if (rc) return rc;
    const uint32_t Lq = c->L * c->q;
    const uint32_t nblk = (Lq + 127) / 128;
    const uint32_t nbw = nblk * (nblk + 1) / 2;
    const uint32_t ksplit = bicount_split(c, nseq);
    Scope s(c, "calcBicounts", (size_t)nseq * c->L);
    if (ksplit > 1) {
        int frc = fill_u32_dev(c, cb->d, 0u, (uint64_t)c->nP * c->q * c->q);
        if (frc) return frc;
    }
    const uint32_t stage = bc_stage_for(c->q);
    auto kern = stage == 16 ? k_bicounts_mfma128<16> : stage == 8 ? k_bicounts_mfma128<8>
                                                                   : k_bicounts_mfma128<4>;
    hipLaunchKernelGGL(kern, dim3(nbw * ksplit), dim3(64), bc_lds_bytes(c->q, stage), c->stream,
                       (const uint8_t*)sb->d, (uint32_t)ld, nseq, c->L, c->q, nblk, ksplit,
                       (uint32_t*)cb->d);
    HIPCHK(c, hipGetLastError());
    cb->version++;
    return MI3_OK;
}

int mi3_bicount_split(mi3_ctx* c, uint32_t nseq, uint32_t* ksplit) {
    if (!c || !ksplit) return MI3_EINVAL;
    *ksplit = bicount_split(c, nseq);
    return MI3_OK;
}

int mi3_bicounts_to_bimarg(mi3_ctx* c, uint32_t nseq) {
    if (!c) return MI3_EINVAL;
    Buf *cb, *bb;
    size_t n = (size_t)c->nP * c->q * c->q;
    int rc = need(c, "bicount", n * 4, &cb);
    if (rc) return rc;
    rc = need(c, "bi", n * 4, &bb);
    if (rc) return rc;
    Scope s(c, "bicounts_to_bimarg", n * 8);
    hipLaunchKernelGGL(k_bicounts_to_bimarg, dim3(grid1(n)), dim3(256), 0, c->stream,
                       (const uint32_t*)cb->d, (float*)bb->d, nseq, (uint64_t)n);
    HIPCHK(c, hipGetLastError());
    bb->version++;
    return MI3_OK;
}

int mi3_weighted_marg(mi3_ctx* c, const char* seqbuf, const char* wbuf, uint32_t nseq) {
    if (!c) return MI3_EINVAL;
    Buf *sb, *wb, *bb;
    int rc = need(c, seqbuf, (size_t)c->L, &sb);
    if (rc) return rc;
    size_t ld = seq_ld(c, sb);
    if (nseq > ld) return fail(c, MI3_EINVAL, "nseq exceeds the sequence buffer");
    if (ld % 16) return fail(c, MI3_EINVAL, "sequence buffer row stride must be a multiple of 16");
    rc = need(c, wbuf, (size_t)nseq * 4, &wb);
    if (rc) return rc;
    size_t qq = (size_t)c->q * c->q;
    rc = need(c, "bi", c->nP * qq * 4, &bb);
    if (rc) return rc;
    const uint64_t nbins = (uint64_t)c->nP * qq;
    if (c->wq_cap < nseq || !c->wq) {
        hipFree(c->wq);
        c->wq = nullptr;
        c->wq_cap = 0;
        HIPCHK(c, hipMalloc(&c->wq, (size_t)nseq * sizeof(uint64_t) + 256));
        c->wq_cap = nseq;
    }
    if (!c->wacc) HIPCHK(c, hipMalloc(&c->wacc, nbins * sizeof(unsigned long long)));
    size_t lds = qq * WM_COLS * sizeof(unsigned long long);
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_weighted_marg),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    Scope s(c, "weightedMarg", (size_t)nseq * (c->L + 8));
    uint32_t nb = grid1(nseq);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(256), 0, c->stream, (const float*)wb->d,
                       nseq, 1, c->red);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->stream, (const double*)c->red,
                       nb, 1, c->red_out);
    hipLaunchKernelGGL(k_weights_fixed, dim3(grid1(nseq ? nseq : 1)), dim3(256), 0, c->stream,
                       (const float*)wb->d, nseq, (const float*)c->red_out, c->wq, c->wbad);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid1(2 * nbins)), dim3(256), 0, c->stream,
                       (uint32_t*)c->wacc, 0u, 2 * nbins);
    hipLaunchKernelGGL(k_weighted_marg, dim3(c->n_units * c->wm_split), dim3(WM_THREADS), lds,
                       c->stream, (const uint8_t*)sb->d, (uint32_t)ld, nseq,
                       (const uint64_t*)c->wq, c->L, c->q, (const uint32_t*)c->wm_units,
                       c->wm_split, c->wacc, 0u);
    hipLaunchKernelGGL(k_fixed_to_float, dim3(grid1(nbins)), dim3(256), 0, c->stream,
                       (const unsigned long long*)c->wacc, (float*)bb->d, nbins,
                       (const float*)c->red_out, (const uint32_t*)c->wbad);
    HIPCHK(c, hipGetLastError());
    bb->version++;
    return MI3_OK;
}

int mi3_weighted_marg_noalloc(mi3_ctx* c, const uint8_t* seq, uint32_t ld, const float* w,
                              uint32_t nseq, float* bi) {
    if (!c) return MI3_EINVAL;
    if (!seq || !w || !bi) return fail(c, MI3_EINVAL, "null device pointer");
    if (nseq > ld) return fail(c, MI3_EINVAL, "nseq exceeds the sequence row stride");
    if (ld % 16) return fail(c, MI3_EINVAL, "sequence buffer row stride must be a multiple of 16");
    if (nseq > MI3_NOALLOC_MAX_SEQ)
        return fail(c, MI3_EINVAL,
                    "nseq = " + std::to_string(nseq) + " exceeds the allocation-free limit of " +
                        std::to_string(MI3_NOALLOC_MAX_SEQ) + " sequences");
    if (!c->scratch) return fail(c, MI3_EHIP, c->scratch_err);
    // scratch regions (this device's copy, by its resolved address)
    uint64_t* wq = reinterpret_cast<uint64_t*>(c->scratch);
    double* red = reinterpret_cast<double*>(c->scratch + NA_RED_OFF);
    float* red_out = reinterpret_cast<float*>(c->scratch + NA_REDOUT_OFF);
    uint32_t* bad = reinterpret_cast<uint32_t*>(c->scratch + NA_BAD_OFF);
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(c->scratch + NA_ACC_OFF);
    const uint64_t qq = (uint64_t)c->q * c->q;
    const size_t lds = qq * WM_COLS * sizeof(unsigned long long);
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_weighted_marg),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // sum of the weights, then fixed-point weights (every kernel below
    // writes what it later reads: nothing in the scratch needs clearing
    // except the accumulator, cleared per run of units)
    uint32_t nb = grid1(nseq);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(256), 0, c->stream, w, nseq, 1, red);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->stream, (const double*)red, nb, 1,
                       red_out);
    hipLaunchKernelGGL(k_weights_fixed, dim3(grid1(nseq ? nseq : 1)), dim3(256), 0, c->stream, w,
                       nseq, (const float*)red_out, wq, bad);
    HIPCHK(c, hipGetLastError());
    // runs of units whose pair blocks fit the accumulator: clear it (a fill
    // kernel, not a memset, so the sequence can be graph-captured), count,
    // and convert the run's bins into their place in `bi`
    uint32_t u0 = 0;
    while (u0 < c->n_units) {
        const uint32_t p_lo = c->unit_pair_lo[u0];
        uint32_t u1 = u0 + 1;
        while (u1 < c->n_units && (uint64_t)(c->unit_pair_hi[u1] - p_lo) * qq <= NA_ACC_BINS) u1++;
        const uint64_t nbins = (uint64_t)(c->unit_pair_hi[u1 - 1] - p_lo) * qq;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid1(2 * nbins)), dim3(256), 0, c->stream,
                           reinterpret_cast<uint32_t*>(acc), 0u, 2 * nbins);
        hipLaunchKernelGGL(k_weighted_marg, dim3((u1 - u0) * c->wm_split), dim3(WM_THREADS), lds,
                           c->stream, seq, ld, nseq, (const uint64_t*)wq, c->L, c->q,
                           (const uint32_t*)(c->wm_units + u0), c->wm_split, acc, p_lo);
        hipLaunchKernelGGL(k_fixed_to_float, dim3(grid1(nbins)), dim3(256), 0, c->stream,
                           (const unsigned long long*)acc, bi + (size_t)p_lo * qq, nbins,
                           (const float*)red_out, (const uint32_t*)bad);
        HIPCHK(c, hipGetLastError());
        u0 = u1;
    }
    return MI3_OK;
}

int mi3_renormalize_bimarg(mi3_ctx* c) {
    if (!c) return MI3_EINVAL;
    Buf* bb;
    size_t qq = (size_t)c->q * c->q;
    int rc = need(c, "bi", c->nP * qq * 4, &bb);
    if (rc) return rc;
    Scope s(c, "renormalize_bimarg", c->nP * qq * 8);
    hipLaunchKernelGGL(k_renormalize, dim3((c->nP + 3) / 4), dim3(256), 0, c->stream,
                       (float*)bb->d, c->nP, (uint32_t)qq);
    HIPCHK(c, hipGetLastError());
    bb->version++;
    return MI3_OK;
}

static int reduce(mi3_ctx* c, const char* buf, uint32_t n, int op, const char* outname) {
    Buf *b, *o;
    int rc = need(c, buf, (size_t)n * 4, &b);
    if (rc) return rc;
    rc = need(c, outname, op == 0 ? 4 : 8, &o);
    if (rc) return rc;
    uint32_t nb = grid1(n);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    if (nb == 0) nb = 1;
    Scope s(c, op == 0 ? "min_buf" : "weight_statistics", (size_t)n * 4);
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(256), 0, c->stream, (const float*)b->d, n,
                       op, c->red);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->stream, (const double*)c->red,
                       nb, op, (float*)o->d);
    HIPCHK(c, hipGetLastError());
    o->version++;
    return MI3_OK;
}

int mi3_min_buf(mi3_ctx* c, const char* buf, uint32_t n) {
    if (!c) return MI3_EINVAL;
    return reduce(c, buf, n, 0, "minout");
}

int mi3_weight_stats(mi3_ctx* c, const char* wbuf, uint32_t n) {
    if (!c) return MI3_EINVAL;
    return reduce(c, wbuf, n, 1, "weightstats");
}

int mi3_dE_to_weights(mi3_ctx* c, const char* Ebuf, const char* wbuf, uint32_t n, float ref) {
    if (!c) return MI3_EINVAL;
    Buf *e, *w;
    int rc = need(c, Ebuf, (size_t)n * 4, &e);
    if (rc) return rc;
    rc = need(c, wbuf, (size_t)n * 4, &w);
    if (rc) return rc;
    Scope s(c, "dE_to_weights", (size_t)n * 8);
    hipLaunchKernelGGL(k_dE_to_weights, dim3(grid1(n)), dim3(256), 0, c->stream,
                       (const float*)e->d, (float*)w->d, ref, n);
    HIPCHK(c, hipGetLastError());
    w->version++;
    return MI3_OK;
}

int mi3_fixed_beta_weights(mi3_ctx* c, const char* Ebuf, const char* wbuf, uint32_t n, float ref,
                           float beta) {
    if (!c) return MI3_EINVAL;
    Buf *e, *w;
    int rc = need(c, Ebuf, (size_t)n * 4, &e);
    if (rc) return rc;
    rc = need(c, wbuf, (size_t)n * 4, &w);
    if (rc) return rc;
    Scope s(c, "fixed_beta_weights", (size_t)n * 8);
    hipLaunchKernelGGL(k_fixed_beta_weights, dim3(grid1(n)), dim3(256), 0, c->stream,
                       (const float*)e->d, (float*)w->d, ref, beta - 1.f, n);
    HIPCHK(c, hipGetLastError());
    w->version++;
    return MI3_OK;
}

int mi3_add_float_bufs(mi3_ctx* c, const char* dst, const char* src, uint32_t n) {
    if (!c) return MI3_EINVAL;
    Buf *d, *s_;
    int rc = need(c, dst, (size_t)n * 4, &d);
    if (rc) return rc;
    rc = need(c, src, (size_t)n * 4, &s_);
    if (rc) return rc;
    Scope s(c, "addbuf", (size_t)n * 12);
    hipLaunchKernelGGL(k_add_floats, dim3(grid1(n)), dim3(256), 0, c->stream, (float*)d->d,
                       (const float*)s_->d, (uint64_t)n);
    HIPCHK(c, hipGetLastError());
    d->version++;
    return MI3_OK;
}

static int update_reg(mi3_ctx* c, const char* Jbuf, float gamma, float pc, float lam, int mode,
                      const char* label) {
    Buf *bt, *bi, *J, *dJ;
    size_t n = (size_t)c->nP * c->q * c->q * 4;
    int rc = need(c, "bi target", n, &bt);
    if (rc) return rc;
    rc = need(c, "bi", n, &bi);
    if (rc) return rc;
    rc = need(c, "J", n, &J);
    if (rc) return rc;
    rc = need(c, Jbuf, n, &dJ);
    if (rc) return rc;
    Scope s(c, label, n * (mode >= 3 ? 5 : 4));
    hipLaunchKernelGGL(k_update_reg, dim3((c->nP + 3) / 4), dim3(256), 0, c->stream,
                       (const float*)bt->d, (const float*)bi->d, (const float*)J->d, (float*)dJ->d,
                       c->nP, c->q, gamma, pc, lam, mode);
    HIPCHK(c, hipGetLastError());
    dJ->version++;
    return MI3_OK;
}

int mi3_update_J(mi3_ctx* c, float gamma, float pc, const char* Jbuf) {
    if (!c) return MI3_EINVAL;
    return update_reg(c, Jbuf ? Jbuf : "dJ", gamma, pc, 0.f, 0, "updateJ");
}

int mi3_reg_l1z(mi3_ctx* c, float gamma, float pc, float lJ) {
    if (!c) return MI3_EINVAL;
    return update_reg(c, "dJ", gamma, pc, lJ, 1, "reg_l1z");
}

int mi3_reg_l2z(mi3_ctx* c, float gamma, float pc, float lJ) {
    if (!c) return MI3_EINVAL;
    return update_reg(c, "dJ", gamma, pc, lJ, 2, "reg_l2z");
}

int mi3_update_J_reg(mi3_ctx* c, float gamma, float pc, int reg, float lJ) {
    if (!c) return MI3_EINVAL;
    if (reg < 0 || reg > 2) return fail(c, MI3_EINVAL, "reg must be 0, 1 (l1z) or 2 (l2z)");
    return update_reg(c, "dJ", gamma, pc, lJ, reg == 0 ? 0 : reg + 2, "updateJ_reg");
}

static int reg_block(mi3_ctx* c, int mode, float gamma, float pc, float p0, float p1, float p2,
                     const char* label) {
    Buf *bi, *J, *dJ, *cr = nullptr;
    size_t n = (size_t)c->nP * c->q * c->q * 4;
    int rc = need(c, "bi", n, &bi);
    if (rc) return rc;
    rc = need(c, "J", n, &J);
    if (rc) return rc;
    rc = need(c, "dJ", n, &dJ);
    if (rc) return rc;
    if (mode == REG_XIJ) {
        rc = need(c, "Creg", (size_t)c->nP * 4, &cr);
        if (rc) return rc;
    }
    Scope s(c, label, n * 4);
    hipLaunchKernelGGL(k_reg_block, dim3((c->nP + 3) / 4), dim3(256), 0, c->stream,
                       (const float*)bi->d, (const float*)J->d, (float*)dJ->d,
                       cr ? (const float*)cr->d : (const float*)nullptr, c->nP, c->q, gamma, pc,
                       p0, p1, p2, mode);
    HIPCHK(c, hipGetLastError());
    dJ->version++;
    return MI3_OK;
}

int mi3_reg_SCADJ(mi3_ctx* c, float gamma, float pc, float r, float a) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_SCADJ, gamma, pc, r, a, 0.f, "reg_SCADJ");
}

int mi3_reg_X(mi3_ctx* c, float gamma, float pc, float lX) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_X, gamma, pc, lX, 0.f, 0.f, "reg_X");
}

int mi3_reg_Xij(mi3_ctx* c, float gamma, float pc) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_XIJ, gamma, pc, 0.f, 0.f, 0.f, "reg_Xij");
}

int mi3_reg_SCADX(mi3_ctx* c, float gamma, float pc, float s, float r, float a) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_SCADX, gamma, pc, s, r, a, "reg_SCADX");
}

int mi3_reg_expX(mi3_ctx* c, float gamma, float pc, float lam) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_EXPX, gamma, pc, lam, 0.f, 0.f, "reg_expX");
}

int mi3_reg_ddE(mi3_ctx* c, float gamma, float pc, float lam) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_DDE, gamma, pc, lam, 0.f, 0.f, "reg_ddE");
}

int mi3_reg_SCADddE(mi3_ctx* c, float gamma, float pc, float lam, float r) {
    if (!c) return MI3_EINVAL;
    return reg_block(c, REG_SCADDDE, gamma, pc, lam, r, 0.f, "reg_SCADddE");
}

int mi3_profile_flush(mi3_ctx* c, const char* path) {
    if (!c) return MI3_EINVAL;
    FILE* f = path ? fopen(path, "a") : nullptr;
    int n = 0;
    size_t keep = 0;
    for (size_t i = 0; i < c->prof.size(); i++) {
        ProfRec& r = c->prof[i];
        if (hipEventQuery(r.b) != hipSuccess) {
            c->prof[keep++] = r;
            continue;
        }
        float ms_a = 0.f, ms_b = 0.f;
        hipEventElapsedTime(&ms_a, c->t0, r.a);
        hipEventElapsedTime(&ms_b, c->t0, r.b);
        if (f)
            fprintf(f, "EVT %s %lld %lld %zu\n", r.name.c_str(), (long long)(ms_a * 1e6),
                    (long long)(ms_b * 1e6), r.bytes);
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
        n++;
    }
    c->prof.resize(keep);
    if (f) fclose(f);
    return n;
}

int mi3_profile_kernel_ms(mi3_ctx* c, const char* name, double* avg_ms, int* count) {
    if (!c || !name || !avg_ms || !count) return MI3_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0;
    int n = 0;
    for (auto& r : c->prof)
        if (r.name == name) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                tot += ms;
                n++;
            }
        }
    *avg_ms = n ? tot / n : 0.0;
    *count = n;
    return MI3_OK;
}

int mi3_profile_reset(mi3_ctx* c) {
    if (!c) return MI3_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& r : c->prof) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    c->prof.clear();
    return MI3_OK;
}

}  // extern "C"
