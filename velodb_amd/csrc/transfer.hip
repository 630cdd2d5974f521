// transfer.hip -- host <-> device wire format: padded, staged, pipelined
// transfers between pageable host memory and device columns.
//
// Reference: StagedTransfer (include/cuda/staged_transfer.hpp:101-225) and
// the padded column moves of ValueVector::to (include/data/value_vector.hpp:140-294).
// The reference copies one chunk at a time (single-threaded memcpy into a
// pinned buffer, DMA, stream synchronize) and some of its moves send the
// exact row count (:151-161).  Here:
//   - every transfer moves `padded_bytes` over the bus: the caller's bytes,
//     then zeros from a pinned zero block (to device) or into a discard
//     buffer (to host), so only the padded size is visible on the link;
//   - a ring of pinned chunks owned by the context, each with an event; the
//     host copy of chunk k+1.. runs while chunk k is on the DMA engine;
//   - the host copies are split over worker threads (one pageable memcpy is
//     far below the PCIe rate).
#include "vo_internal.hpp"

#include <condition_variable>
#include <functional>
#include <thread>

namespace vo {
namespace {

constexpr size_t kChunk = 32ull << 20; // bytes per staging chunk
constexpr int kRing = 4;                // chunks in flight
constexpr int kWorkers = 8;             // host copy threads (the caller is one more)

// A fixed pool of threads that copy slices of one memcpy in parallel.
class CopyPool {
public:
    CopyPool()
    {
        for (int i = 0; i < kWorkers; ++i) th_.emplace_back([this, i] { loop(i); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    // memcpy(dst, src, bytes) on all threads; returns when done
    void copy(void *dst, const void *src, size_t bytes)
    {
        if (bytes < (1u << 20)) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            bytes_ = bytes;
            pending_ = kWorkers;
            ++gen_;
        }
        cv_.notify_all();
        slice(kWorkers); // the caller takes the last slice
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
    }

private:
    void slice(int part)
    {
        const size_t per = ((bytes_ / (kWorkers + 1)) + 4095) & ~(size_t)4095;
        const size_t lo = std::min(bytes_, per * part);
        const size_t hi = part == kWorkers ? bytes_ : std::min(bytes_, lo + per);
        if (hi > lo) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void loop(int part)
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
            }
            slice(part);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    bool quit_ = false;
    uint64_t gen_ = 0;
    int pending_ = 0;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
};

struct Staging {
    char *chunk[kRing] = {};
    hipEvent_t ev[kRing] = {};
    bool busy[kRing] = {};
    char *zeros = nullptr;   // kChunk zero bytes, the pad rows of uploads
    char *discard = nullptr; // kChunk bytes receiving the pad rows of downloads
    hipStream_t stream = nullptr;
    CopyPool pool;
    uint64_t bytes_moved = 0; // bytes over the link, last transfer
    ~Staging()
    {
        for (int i = 0; i < kRing; ++i) {
            if (chunk[i]) (void)hipHostFree(chunk[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
        if (zeros) (void)hipHostFree(zeros);
        if (discard) (void)hipHostFree(discard);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

Staging &staging(vo_ctx *ctx)
{
    if (!ctx->staging) {
        auto *s = new Staging();
        ctx->staging = s;
        for (int i = 0; i < kRing; ++i) {
            VO_HIP(hipHostMalloc((void **)&s->chunk[i], kChunk, hipHostMallocDefault));
            VO_HIP(hipEventCreateWithFlags(&s->ev[i], hipEventDisableTiming));
        }
        VO_HIP(hipHostMalloc((void **)&s->zeros, kChunk, hipHostMallocDefault));
        std::memset(s->zeros, 0, kChunk);
        VO_HIP(hipHostMalloc((void **)&s->discard, kChunk, hipHostMallocDefault));
        // transfers run on their own stream, ordered after the context's
        // stream at start and before it at the end
        VO_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    return *(Staging *)ctx->staging;
}

// make `a` wait for the work enqueued so far on `b`
void order_after(hipStream_t a, hipStream_t b)
{
    hipEvent_t e;
    VO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    VO_HIP(hipEventRecord(e, b));
    VO_HIP(hipStreamWaitEvent(a, e, 0));
    VO_HIP(hipEventDestroy(e));
}

} // namespace

void staging_free(vo_ctx *ctx)
{
    if (!ctx->staging) return;
    delete (Staging *)ctx->staging;
    ctx->staging = nullptr;
}

uint64_t last_transfer_bytes(vo_ctx *ctx) { return ctx->staging ? ((Staging *)ctx->staging)->bytes_moved : 0; }

void op_transfer_to_device(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint64_t padded_bytes)
{
    VO_CHECK(padded_bytes >= bytes, VO_ERR_INVALID, "padded_bytes below bytes");
    VO_CHECK((dst && src) || bytes == 0, VO_ERR_INVALID, "null buffer");
    VO_CHECK(dst || padded_bytes == 0, VO_ERR_INVALID, "null buffer");
    Staging &s = staging(ctx);
    order_after(s.stream, ctx->stream);
    s.bytes_moved = 0;
    uint64_t off = 0;
    int k = 0;
    while (off < bytes) {
        const size_t n = (size_t)std::min<uint64_t>(kChunk, bytes - off);
        if (s.busy[k]) VO_HIP(hipEventSynchronize(s.ev[k])); // the chunk's previous DMA is done
        s.pool.copy(s.chunk[k], (const char *)src + off, n);
        VO_HIP(hipMemcpyAsync((char *)dst + off, s.chunk[k], n, hipMemcpyHostToDevice, s.stream));
        VO_HIP(hipEventRecord(s.ev[k], s.stream));
        s.busy[k] = true;
        off += n;
        s.bytes_moved += n;
        k = (k + 1) % kRing;
    }
    while (off < padded_bytes) { // zeros over the link for the pad rows
        const size_t n = (size_t)std::min<uint64_t>(kChunk, padded_bytes - off);
        VO_HIP(hipMemcpyAsync((char *)dst + off, s.zeros, n, hipMemcpyHostToDevice, s.stream));
        off += n;
        s.bytes_moved += n;
    }
    order_after(ctx->stream, s.stream);
    VO_HIP(hipStreamSynchronize(s.stream));
    for (int i = 0; i < kRing; ++i) s.busy[i] = false;
}

void op_transfer_to_host(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint64_t padded_bytes)
{
    VO_CHECK(padded_bytes >= bytes, VO_ERR_INVALID, "padded_bytes below bytes");
    VO_CHECK((dst && src) || bytes == 0, VO_ERR_INVALID, "null buffer");
    VO_CHECK(src || padded_bytes == 0, VO_ERR_INVALID, "null buffer");
    Staging &s = staging(ctx);
    order_after(s.stream, ctx->stream);
    s.bytes_moved = 0;
    // chunk c of the caller's bytes goes through ring slot c % kRing: DMAs
    // run up to kRing chunks ahead of the host copies
    const uint64_t nchunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](uint64_t c) {
        const uint64_t off = c * kChunk;
        const size_t n = (size_t)std::min<uint64_t>(kChunk, bytes - off);
        const int k = (int)(c % kRing);
        VO_HIP(hipMemcpyAsync(s.chunk[k], (const char *)src + off, n, hipMemcpyDeviceToHost, s.stream));
        VO_HIP(hipEventRecord(s.ev[k], s.stream));
    };
    for (uint64_t c = 0; c < std::min<uint64_t>(nchunks, kRing); ++c) issue(c);
    for (uint64_t c = 0; c < nchunks; ++c) {
        const int k = (int)(c % kRing);
        const uint64_t off = c * kChunk;
        const size_t n = (size_t)std::min<uint64_t>(kChunk, bytes - off);
        VO_HIP(hipEventSynchronize(s.ev[k]));
        s.pool.copy((char *)dst + off, s.chunk[k], n);
        s.bytes_moved += n;
        if (c + kRing < nchunks) issue(c + kRing);
    }
    for (uint64_t off = bytes; off < padded_bytes;) { // pad rows over the link, discarded
        const size_t n = (size_t)std::min<uint64_t>(kChunk, padded_bytes - off);
        VO_HIP(hipMemcpyAsync(s.discard, (const char *)src + off, n, hipMemcpyDeviceToHost, s.stream));
        off += n;
        s.bytes_moved += n;
    }
    VO_HIP(hipStreamSynchronize(s.stream));
}

} // namespace vo
