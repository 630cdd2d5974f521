// resident.hpp -- recorded command programs for the resident queue kernel.
//
// In persistent mode an operator is not launched: its host driver runs as in
// launch mode, but every launch(), fill and copy it issues is appended to a
// program of steps.  A step names the device body of the kernel it replaces,
// its grid (virtual blocks), its dynamic LDS and its arguments packed with
// the kernel's own parameter layout.  The resident kernel (queue.hip) runs a
// program step by step: each step's virtual blocks are spread over the
// resident blocks, with a grid barrier between steps.  The same host code and
// the same device bodies run in both modes, so both give the same bytes.
//
// Every kernel with a resident body is listed in one X-macro table per source
// file (VO_*_BODIES): the host side maps the kernel's address to the entry's
// id, the resident kernel switches on the id and calls the body.
#pragma once

// Included at the end of vo_internal.hpp (needs vo_ctx and LaunchInfo).
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

namespace vo {

constexpr uint32_t kStepBlob = 240;      // packed kernel arguments in a step's first slot
constexpr uint32_t kStepMaxSlots = 4;    // a step with larger arguments runs on into the next slots
constexpr uint32_t kStepMaxBlob = kStepBlob + (kStepMaxSlots - 1) * 256;
constexpr uint32_t kResidentThreads = 256;
constexpr uint32_t kResidentLds = 48u << 10; // dynamic LDS of one resident block (largest body: 43.5 KiB)

// One 256-byte program slot.  A step whose arguments exceed kStepBlob bytes
// takes `slots` consecutive slots: its blob continues over the following
// slots' bytes.
struct alignas(16) Step {
    uint32_t id;    // StepId or a body id from the tables
    uint32_t grid;  // virtual blocks
    uint32_t lds;   // dynamic LDS bytes the body reads from its `lds` argument
    uint16_t aux;   // fill: byte value (low 8 bits); kStepNoBarrier
    uint16_t slots; // program slots this step takes (>= 1)
    char blob[kStepBlob];
};
static_assert(sizeof(Step) == 256, "one program slot is 256 bytes");

// Step::aux flag: the next step does not read what this one writes, so the
// lane runs on into it without a barrier.
constexpr uint16_t kStepNoBarrier = 0x8000;

// Built-in steps (ids below kFirstBody).
enum StepId : uint32_t {
    kStepFill = 1, // blob: {void *dst; uint64_t bytes}, aux = byte value
    kStepCopy = 2, // blob: {void *dst; const void *src; uint64_t bytes}; src may be device or host-mapped memory
    kFirstBody = 16,
};

struct FillArgs {
    void *dst;
    uint64_t bytes;
};
struct CopyArgs {
    void *dst;
    const void *src;
    uint64_t bytes;
};

// Parameter layout of a kernel type void(P...): each parameter at the next
// offset aligned for its type, as in a struct.
template <class F>
struct KSig;
template <class... P>
struct KSig<void(P...)> {
    static constexpr size_t n = sizeof...(P);
    template <size_t I>
    using T = std::tuple_element_t<I, std::tuple<P...>>;
    __host__ __device__ static constexpr size_t off(size_t i)
    {
        constexpr size_t sz[] = {sizeof(P)..., 0};
        constexpr size_t al[] = {alignof(P)..., 1};
        size_t o = 0;
        for (size_t k = 0;; ++k) {
            o = (o + al[k] - 1) / al[k] * al[k];
            if (k == i) return o;
            o += sz[k];
        }
    }
    __host__ __device__ static constexpr size_t bytes() { return n ? off(n - 1) + sizeof(T<n - 1>) : 0; }
};

template <class K>
using KSigOf = KSig<std::remove_pointer_t<std::remove_cv_t<std::remove_reference_t<K>>>>;

// Host: pack `args` (converted to the kernel's parameter types) into blob.
template <class S, size_t... I, class... A>
inline void pack_args_i(char *blob, std::index_sequence<I...>, A... a)
{
    (void)blob;
    (..., [&] {
        const typename S::template T<I> v = a;
        std::memcpy(blob + S::off(I), &v, sizeof(v));
    }());
}

template <class K, class... A>
inline void pack_args(char *blob, A... a)
{
    using S = KSigOf<K>;
    static_assert(S::n == sizeof...(A), "argument count of the kernel");
    static_assert(S::bytes() <= kStepMaxBlob, "kernel arguments exceed the largest step");
    std::memset(blob, 0, kStepMaxBlob);
    pack_args_i<S>(blob, std::make_index_sequence<S::n>{}, a...);
}

// Device: call f with the kernel's parameters read back from blob.
template <class S, class F, size_t... I>
__device__ __forceinline__ void call_blob_i(const char *blob, F &&f, std::index_sequence<I...>)
{
    f(*reinterpret_cast<const typename S::template T<I> *>(blob + S::off(I))...);
}

template <class K, class F>
__device__ __forceinline__ void call_blob(const char *blob, F &&f)
{
    using S = KSigOf<K>;
    call_blob_i<S>(blob, f, std::make_index_sequence<S::n>{});
}

// ------------------------------------------------------------- recording

// A command's program while its host driver runs (vo_ctx::rec).  Data the
// driver uploads from host memory is kept beside the steps and copied by a
// step from the command's host-mapped program area.
struct Recorder {
    uint32_t lane = 0;                   // queue lane the command runs on
    std::vector<Step> steps;
    std::vector<char> data;              // uploaded host bytes
    std::vector<std::pair<size_t, size_t>> data_refs; // (step, offset into data) to relocate
};

// One registry entry: a kernel's host address -> its resident body id and
// the static LDS its body carves from the `lds` argument beyond the launch's
// own dynamic LDS.
struct BodyInfo {
    uint32_t id;
    uint32_t lds_extra;
};

void register_bodies(const std::pair<const void *, BodyInfo> *entries, size_t n);
const BodyInfo *find_body(const void *kernel);

template <typename K, typename... Args>
inline void record_launch(vo_ctx *ctx, const LaunchInfo &li, K kernel, Args... args)
{
    const BodyInfo *bi = find_body((const void *)kernel);
    VO_CHECK(bi != nullptr, VO_ERR_UNSUPPORTED,
             std::string("resident queue: kernel ") + li.family + " has no resident body");
    VO_CHECK(li.grid.y == 1 && li.grid.z == 1, VO_ERR_INVALID, "resident steps take one-dimensional grids");
    const uint32_t lds = li.lds + bi->lds_extra;
    VO_CHECK(lds <= kResidentLds, VO_ERR_INVALID, std::string("resident queue: LDS of ") + li.family);
    constexpr size_t bytes = KSigOf<K>::bytes();
    constexpr uint32_t slots = bytes <= kStepBlob ? 1 : 1 + (uint32_t)((bytes - kStepBlob + 255) / 256);
    Step s[kStepMaxSlots] = {};
    s[0].id = bi->id;
    s[0].grid = li.grid.x;
    s[0].lds = lds;
    s[0].slots = (uint16_t)slots;
    char blob[kStepMaxBlob];
    pack_args<K>(blob, args...);
    std::memcpy(s[0].blob, blob, bytes); // runs on over the following slots
    ctx->rec->steps.insert(ctx->rec->steps.end(), s, s + slots);
}

// Launches `kernel` on the context stream, or appends it as a step while the
// context records a command program.
template <typename K, typename... Args>
inline void launch(vo_ctx *ctx, const LaunchInfo &li, K kernel, Args... args)
{
    if (ctx->rec) {
        record_launch(ctx, li, kernel, args...);
        return;
    }
    hipEvent_t ev = nullptr;
    begin_launch(ctx, li, &ev);
    hipLaunchKernelGGL(kernel, li.grid, li.block, li.lds, ctx->stream, args...);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw Error{VO_ERR_HIP, std::string("launch ") + li.family + ": " + hipGetErrorString(e)};
    end_launch(ctx, ev);
}

// Fill / copy on the context stream, or as steps while recording.
void fill_async(vo_ctx *ctx, void *dst, int value, uint64_t bytes);
void copy_async(vo_ctx *ctx, void *dst, const void *src, uint64_t bytes); // device to device
void upload_async(vo_ctx *ctx, void *dst, const void *host_src, uint64_t bytes); // host to device
// device to pinned (host-mapped) memory; while recording a copy step writes it
void to_host_async(vo_ctx *ctx, void *pinned_dst, const void *src, uint64_t bytes);

// Independent pieces of one command: f(i) for i < n, where no piece reads
// what another writes.  In launch mode they run in order on the context
// stream.  While recording, each piece's steps go to a recorder of their own
// and are merged phase by phase into the command's program: phase k holds
// step k of every piece, with no barrier between the steps of one phase and a
// barrier after it, so the pieces share the lane's blocks.
void run_tracks(vo_ctx *ctx, size_t n, const std::function<void(size_t)> &f);

// The recorded program of one command (persistent mode): runs the command's
// host driver with a recorder attached.
void record_command(vo_ctx *ctx, const vo_command &cmd, Recorder &rec);

} // namespace vo
