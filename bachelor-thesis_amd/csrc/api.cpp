// api.cpp -- library state, error reporting, memory/stream helpers, seeded
// input generation and hex text I/O.
#include "common.h"
#include <cstdarg>
#include <cstring>
#include <cctype>
#include <mutex>
#include <vector>

namespace mgx {

static thread_local std::string g_err;
static std::mutex g_mu;
static bool g_init = false;
static std::vector<unsigned *> g_errword;  // per device

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int fail(int status, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

int noalloc_scratch_init(int dev);   // crt.hip

int ensure_init()
{
    std::lock_guard<std::mutex> lk(g_mu);
    int dev = 0;
    MGX_HIP(hipGetDevice(&dev));
    MGX_TRY(noalloc_scratch_init(dev));
    if ((int)g_errword.size() <= dev) g_errword.resize(dev + 1, nullptr);
    if (!g_errword[dev]) {
        unsigned *w = nullptr;
        MGX_HIP(hipMalloc(&w, 256));
        MGX_HIP(hipMemset(w, 0, 256));
        g_errword[dev] = w;
    }
    g_init = true;
    return MGX_OK;
}

unsigned *device_error_word()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if ((int)g_errword.size() <= dev) return nullptr;
    return g_errword[dev];
}

int check_device_error(hipStream_t s)
{
    unsigned *w = device_error_word();
    if (!w) return fail(MGX_ERR_ARG, "library not initialised");
    unsigned h = 0;
    MGX_HIP(hipMemcpyAsync(&h, w, sizeof h, hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    if (h) {
        MGX_HIP(hipMemsetAsync(w, 0, sizeof h, s));
        MGX_HIP(hipStreamSynchronize(s));
        if (h & DEV_ERR_CAPACITY) return fail(MGX_ERR_CAPACITY, "an output entry needs more limbs than cap");
        if (h & DEV_ERR_INPUT) return fail(MGX_ERR_SHAPE, "an input entry exceeds the plan's bit bound");
        if (h & DEV_ERR_WIDTH)
            return fail(MGX_ERR_UNSUPPORTED, "an operand is wider than the plain route's 512-word kernel");
        return fail(MGX_ERR_HIP, "device error word 0x%x", h);
    }
    return MGX_OK;
}

// Scratch blocks are cached per device for the life of the process (or until
// mgx_scratch_release): a product's residues, Y planes and column sums take
// over a gigabyte, and mapping that afresh costs the better part of a
// millisecond per call.  A block freed on a stream carries an event recorded
// there; whichever stream takes it next waits on that event, so reuse stays
// stream-ordered without a host synchronization.  A request takes the smallest
// idle block that holds it and is at most about twice its size; otherwise a
// new block is mapped, and idle blocks that the new one supersedes (large ones
// smaller than it) are unmapped.  Work being captured into a graph gets
// graph-owned memory from the stream-ordered allocator instead.
namespace {
struct ScratchBlock {
    void *p;
    size_t size;
    hipEvent_t ev;   // recorded where the block was last released
    bool busy;
};
std::mutex g_scratch_mu;
std::vector<std::vector<ScratchBlock>> g_scratch;   // per device
constexpr size_t SCRATCH_SLACK = 4u << 20;
constexpr size_t SCRATCH_MAX_BLOCKS = 64;

void unmap_block(ScratchBlock &b)
{
    (void)hipEventSynchronize(b.ev);
    (void)hipEventDestroy(b.ev);
    (void)hipFree(b.p);
}
}  // namespace

int scratch_alloc(void **p, size_t bytes, hipStream_t s)
{
    bytes = (size_t)round_up((int64_t)std::max<size_t>(bytes, 1), 256);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    MGX_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) {
        hipError_t e = hipMallocAsync(p, bytes, s);
        if (e != hipSuccess) return fail(MGX_ERR_OOM, "hipMallocAsync(%zu): %s", bytes, hipGetErrorString(e));
        return MGX_OK;
    }
    int dev = 0;
    MGX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    auto &blocks = g_scratch[dev];
    ScratchBlock *best = nullptr;
    for (auto &b : blocks)
        if (!b.busy && b.size >= bytes && b.size <= 2 * bytes + SCRATCH_SLACK && (!best || b.size < best->size))
            best = &b;
    if (best) {
        MGX_HIP(hipStreamWaitEvent(s, best->ev, 0));
        best->busy = true;
        *p = best->p;
        return MGX_OK;
    }
    for (size_t i = blocks.size(); i-- > 0;) {
        ScratchBlock &b = blocks[i];
        bool superseded = b.size > SCRATCH_SLACK && b.size < bytes;
        if (!b.busy && (superseded || blocks.size() >= SCRATCH_MAX_BLOCKS)) {
            unmap_block(b);
            blocks.erase(blocks.begin() + (ptrdiff_t)i);
        }
    }
    ScratchBlock nb{nullptr, bytes, nullptr, true};
    hipError_t e = hipMalloc(&nb.p, bytes);
    if (e != hipSuccess) return fail(MGX_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipEventCreateWithFlags(&nb.ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipFree(nb.p);
        return fail(MGX_ERR_HIP, "hipEventCreateWithFlags: %s", hipGetErrorString(e));
    }
    blocks.push_back(nb);
    *p = nb.p;
    return MGX_OK;
}

int scratch_free(void *p, hipStream_t s)
{
    if (!p) return MGX_OK;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        for (auto &blocks : g_scratch)
            for (auto &b : blocks)
                if (b.p == p && b.busy) {
                    MGX_HIP(hipEventRecord(b.ev, s));
                    b.busy = false;
                    return MGX_OK;
                }
    }
    MGX_HIP(hipFreeAsync(p, s));   // graph-owned memory from a captured call
    return MGX_OK;
}

int scratch_release_device(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) return MGX_OK;
    auto &blocks = g_scratch[dev];
    for (size_t i = blocks.size(); i-- > 0;)
        if (!blocks[i].busy) {
            unmap_block(blocks[i]);
            blocks.erase(blocks.begin() + (ptrdiff_t)i);
        }
    return MGX_OK;
}

static int64_t g_params[8] = {256, 2, -1, 0, 0, 0, 0, 0};
int64_t param(int which) { return (which >= 0 && which < 8) ? g_params[which] : 0; }

}  // namespace mgx

using namespace mgx;

extern "C" {

int mgx_init(int device)
{
    if (device >= 0) MGX_HIP(hipSetDevice(device));
    return ensure_init();
}

const char *mgx_last_error(void) { return g_err.c_str(); }
const char *mgx_version(void) { return "mgx 0.1 gfx950"; }

int mgx_set_param(int which, int64_t value)
{
    if (which < 0 || which > MGX_PARAM_RECON) return fail(MGX_ERR_ARG, "unknown parameter %d", which);
    if (which == MGX_PARAM_REDUCE && value != 0 && value != 1) return fail(MGX_ERR_ARG, "reduce kernel must be 0 or 1");
    if (which == MGX_PARAM_RECON && (value < 0 || value > 2)) return fail(MGX_ERR_ARG, "reconstruct mode must be 0, 1 or 2");
    if (which == MGX_PARAM_GEMM_ROWS && value != 0 && value != 64 && value != 128)
        return fail(MGX_ERR_ARG, "GEMM tile rows must be 0, 64 or 128");
    if (which == MGX_PARAM_TRANSFER && value < 0) return fail(MGX_ERR_ARG, "transfer size must be >= 0");
    if (which == MGX_PARAM_CRT_SLICES && value != 0 && value != 2 && value != 4)
        return fail(MGX_ERR_ARG, "CRT slice count must be 0, 2 or 4");
    if (which == MGX_PARAM_LEAF && value < 1) return fail(MGX_ERR_ARG, "leaf size must be >= 1");
    g_params[which] = value;
    return MGX_OK;
}
int64_t mgx_get_param(int which) { return param(which); }

int mgx_malloc(void **p, size_t bytes)
{
    MGX_TRY(ensure_init());
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(MGX_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return MGX_OK;
}
int mgx_free(void *p)
{
    if (p) MGX_HIP(hipFree(p));
    return MGX_OK;
}
int mgx_scratch_release(void)
{
    int dev = 0;
    MGX_HIP(hipGetDevice(&dev));
    return scratch_release_device(dev);
}
int mgx_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream)
{
    MGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    MGX_HIP(hipStreamSynchronize(as_stream(stream)));
    return MGX_OK;
}
int mgx_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream)
{
    MGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    MGX_HIP(hipStreamSynchronize(as_stream(stream)));
    return MGX_OK;
}
int mgx_memset(void *dst, int value, size_t bytes, void *stream)
{
    MGX_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    return MGX_OK;
}
int mgx_stream_create(void **stream)
{
    MGX_TRY(ensure_init());
    hipStream_t s;
    MGX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return MGX_OK;
}
int mgx_stream_destroy(void *stream)
{
    if (stream) MGX_HIP(hipStreamDestroy(as_stream(stream)));
    return MGX_OK;
}
int mgx_stream_sync(void *stream)
{
    MGX_HIP(hipStreamSynchronize(as_stream(stream)));
    return MGX_OK;
}
int mgx_device_sync(void)
{
    MGX_HIP(hipDeviceSynchronize());
    return MGX_OK;
}
int mgx_event_create(void **ev)
{
    hipEvent_t e;
    MGX_HIP(hipEventCreate(&e));
    *ev = e;
    return MGX_OK;
}
int mgx_event_destroy(void *ev)
{
    if (ev) MGX_HIP(hipEventDestroy((hipEvent_t)ev));
    return MGX_OK;
}
int mgx_event_record(void *ev, void *stream)
{
    MGX_HIP(hipEventRecord((hipEvent_t)ev, as_stream(stream)));
    return MGX_OK;
}
int mgx_event_elapsed(float *ms, void *start, void *stop)
{
    MGX_HIP(hipEventSynchronize((hipEvent_t)stop));
    MGX_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return MGX_OK;
}

// ---- seeded generator: SplitMix64 seeding xoshiro256** (Blackman & Vigna) ----
static uint64_t splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
struct xoshiro {
    uint64_t s[4];
    explicit xoshiro(uint64_t seed)
    {
        for (int i = 0; i < 4; i++) s[i] = splitmix(seed);
    }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
};

// Per entry, in row-major order: ceil(nbits/64) limbs low to high (the top
// one masked to nbits % 64 bits), then one word whose low bit is the sign.
int mgx_random_matrix(mgx_limb *host, int64_t W, int64_t H, int64_t rowstride, int64_t cap, int64_t nbits,
                      uint64_t seed)
{
    if (nbits < 0) return fail(MGX_ERR_ARG, "nbits < 0");
    int64_t nl = (nbits + 63) / 64;
    if (nl > cap) return fail(MGX_ERR_CAPACITY, "%lld-bit entries need %lld limbs > cap %lld", (long long)nbits,
                              (long long)nl, (long long)cap);
    xoshiro rng(seed);
    for (int64_t y = 0; y < H; y++)
        for (int64_t x = 0; x < W; x++) {
            mgx_limb *e = host + (y * rowstride + x) * (cap + 1);
            for (int64_t i = 0; i < nl; i++) e[i] = rng.next();
            if (nbits % 64) e[nl - 1] &= (1ull << (nbits % 64)) - 1;
            for (int64_t i = nl; i < cap; i++) e[i] = 0;
            int64_t sz = nl;
            while (sz > 0 && e[sz - 1] == 0) sz--;
            bool neg = rng.next() & 1;
            ((int64_t *)e)[cap] = neg ? -sz : sz;
        }
    return MGX_OK;
}

// The text form of an entry (the format mat-gpu prints, mat-gpu/mpz.cu:85-104,
// 182-204): the magnitude in lower-case hexadecimal without leading zeros,
// preceded by '-' when the value is negative; zero is "0".  The magnitude
// is limbs 0 .. |size|-1 (zero limbs at the top are allowed).  Returns the
// length written (NUL excluded), or -1 when buf is too small or |size| > cap.
int64_t mgx_entry_to_hex(char *buf, int64_t buflen, const mgx_limb *e, int64_t cap)
{
    static const char digits[] = "0123456789abcdef";
    const int64_t size = ((const int64_t *)e)[cap];
    int64_t top = size < 0 ? -size : size;
    if (top > cap || buflen < 2) return -1;
    while (top > 0 && e[top - 1] == 0) top--;
    if (top == 0) {
        buf[0] = '0';
        buf[1] = 0;
        return 1;
    }
    // digits of the top limb, then 16 per lower limb
    int lead = 16;
    while (lead > 1 && ((e[top - 1] >> (4 * (lead - 1))) & 15) == 0) lead--;
    const int64_t len = (size < 0) + lead + 16 * (top - 1);
    if (len + 1 > buflen) return -1;
    char *out = buf;
    if (size < 0) *out++ = '-';
    for (int64_t i = top - 1; i >= 0; i--)
        for (int k = (i == top - 1 ? lead : 16) - 1; k >= 0; k--) *out++ = digits[(e[i] >> (4 * k)) & 15];
    *out = 0;
    return len;
}

// mat-gpu/mpz.cu:106-146 (inp_str): leading space, '-', optional "0x", hex.
int mgx_entry_from_hex(mgx_limb *e, int64_t cap, const char *str)
{
    while (isspace((unsigned char)*str)) str++;
    bool neg = false;
    if (*str == '-') { neg = true; str++; }
    if (str[0] == '0' && str[1] == 'x') str += 2;
    int64_t nch = 0;
    while (str[nch] && !isspace((unsigned char)str[nch])) nch++;
    if (nch > cap * 16) return fail(MGX_ERR_CAPACITY, "hex text of %lld digits exceeds cap", (long long)nch);
    memset(e, 0, (size_t)cap * sizeof(mgx_limb));
    for (int64_t i = 0; i < nch; i++) {
        int c = tolower((unsigned char)str[nch - 1 - i]), d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else return fail(MGX_ERR_PARSE, "invalid character '%c' in hex entry", c);
        e[i / 16] |= (mgx_limb)d << (4 * (i % 16));
    }
    int64_t sz = (nch + 15) / 16;
    ((int64_t *)e)[cap] = neg ? -sz : sz;
    return MGX_OK;
}

}  // extern "C"
