// Host emulation of the small HIP subset used by the generated kernels, for CPU tests of the
// generator: each work-item is a ucontext fiber, __syncthreads() yields to a scheduler that runs
// all fibers of a work-group to the next barrier, work-groups run one after another.
// Test infrastructure only; the library never runs kernels this way.
#ifndef BBFFT_TEST_HIP_EMU_HPP
#define BBFFT_TEST_HIP_EMU_HPP

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <sys/mman.h>
#include <ucontext.h>
#include <vector>

// Vector values start as NaN so that a kernel reading a component it never wrote (on the GPU:
// whatever the register held) produces NaN in the emulated output instead of a benign zero.
struct float2 {
    float x = __builtin_nanf(""), y = __builtin_nanf("");
};
struct double2 {
    double x = __builtin_nan(""), y = __builtin_nan("");
};
struct emu_dim3 {
    unsigned x = 0, y = 0, z = 0;
};

#define __global__
#define __device__
#define __forceinline__ inline
#define __constant__
#define __shared__ static
#define __launch_bounds__(x)
#define __builtin_nontemporal_load(p) (*(p))
// direct global-to-LDS load: the work-item's word lands at the wave's LDS base + its lane
#define BBFFT_LDS_DMA4(src, dst) (((float*)(dst))[threadIdx.x % 64u] = *(const float*)(src))
// 16-byte form: the work-item's four words land at the wave's LDS base + 16 bytes x lane
#define BBFFT_LDS_DMA16(src, dst)                                                               \
    do {                                                                                        \
        for (int q_ = 0; q_ < 4; ++q_)                                                          \
            ((float *)(dst))[4u * (threadIdx.x % 64u) + q_] = ((const float *)(src))[q_];      \
    } while (0)
// the same with the non-temporal cache policy (no effect on the values)
#define BBFFT_LDS_DMA4_NT(src, dst) BBFFT_LDS_DMA4(src, dst)
#define BBFFT_LDS_DMA16_NT(src, dst) BBFFT_LDS_DMA16(src, dst)
#define BBFFT_WAIT_ALL() ((void)0)
#define BBFFT_LDS_SYNC() __syncthreads()
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

// Race checking (built with -fsanitize=thread): every work-item runs on a ThreadSanitizer
// fiber of its own, fiber switches create no ordering, and barrier number e of a work-group
// orders everything its work-items did before it before everything they do after it
// (release on a token of that barrier, acquire when leaving it). Work-groups are ordered one
// after another (the emulator reuses one copy of the LDS arrays for all of them). So
// ThreadSanitizer reports exactly the accesses of two work-items of one group to the same LDS
// or global address, at least one a write, with no barrier between them. The scheduler's own
// bookkeeping uses relaxed atomics, which order nothing and are not reported.
#if defined(__SANITIZE_THREAD__)
#define BBFFT_EMU_TSAN 1
#elif defined(__has_feature)
#if __has_feature(thread_sanitizer)
#define BBFFT_EMU_TSAN 1
#endif
#endif
#ifdef BBFFT_EMU_TSAN
#include <sanitizer/tsan_interface.h>
#endif

namespace emu {
inline emu_dim3 blockIdx_, gridDim_, blockDim_;
inline ucontext_t sched_ctx;
inline std::vector<ucontext_t> fiber_ctx;
inline std::vector<int> fiber_state; // 0 runnable, 1 at barrier, 2 done
inline unsigned current = 0;
inline std::size_t barrier_count = 0;
#ifdef BBFFT_EMU_TSAN
inline std::vector<void *> tsan_fiber;       // one per work-item
inline void *tsan_sched = nullptr;           // the scheduler's (the launching thread's) fiber
inline std::vector<std::size_t> fiber_epoch; // barriers passed by each work-item in its group
inline std::vector<char> epoch_token(1 << 16); // release/acquire token of barrier e (fixed:
                                                 // tokens must not move while in use)
inline char group_token;                     // orders consecutive work-groups
#endif

inline int state(unsigned i) { return __atomic_load_n(&fiber_state[i], __ATOMIC_RELAXED); }
inline void set_state(unsigned i, int v) { __atomic_store_n(&fiber_state[i], v, __ATOMIC_RELAXED); }
inline unsigned cur() { return __atomic_load_n(&current, __ATOMIC_RELAXED); }
inline emu_dim3 thread_idx() {
    emu_dim3 d;
    d.x = cur();
    return d;
}
} // namespace emu

#define threadIdx (emu::thread_idx())
#define blockIdx (emu::blockIdx_)
#define gridDim (emu::gridDim_)
#define blockDim (emu::blockDim_)

inline void __syncthreads() {
    unsigned me = emu::cur();
#ifdef BBFFT_EMU_TSAN
    std::size_t const e = __atomic_fetch_add(&emu::fiber_epoch[me], 1, __ATOMIC_RELAXED);
    __tsan_release(&emu::epoch_token[e]);
    __tsan_switch_to_fiber(emu::tsan_sched, __tsan_switch_to_fiber_no_sync);
#endif
    emu::set_state(me, 1);
    swapcontext(&emu::fiber_ctx[me], &emu::sched_ctx);
#ifdef BBFFT_EMU_TSAN
    __tsan_acquire(&emu::epoch_token[e]);
#endif
}

namespace emu {

#ifdef BBFFT_EMU_TSAN
#define BBFFT_EMU_NO_TSAN __attribute__((no_sanitize("thread")))
#else
#define BBFFT_EMU_NO_TSAN
#endif

/// Entry of a work-item's fiber. Under the race checker it does not return: its function
/// exit would be recorded after the switch back to the scheduler's fiber and pop the
/// scheduler's shadow call stack (compilers record entry and exit even for functions whose
/// memory accesses are not instrumented), so it jumps back to the scheduler instead.
template <typename Body> struct fiber_entry {
    static inline Body *body = nullptr;
    BBFFT_EMU_NO_TSAN static void run() {
#ifdef BBFFT_EMU_TSAN
        __tsan_acquire(&group_token);
#endif
        (*body)();
        set_state(cur(), 2);
#ifdef BBFFT_EMU_TSAN
        __tsan_release(&group_token); // before the next group's work-items
        __tsan_switch_to_fiber(tsan_sched, __tsan_switch_to_fiber_no_sync);
        setcontext(&sched_ctx);
#endif
    }
};

/// Runs body() for `block` work-items of each of `grid` work-groups. Returns 0 on success,
/// 1 if work-items of one group disagree on the number of barriers (divergent barrier).
template <typename Body> int launch(unsigned grid, unsigned block, Body body) {
    static std::size_t const stack_size = 512 * 1024;
    fiber_entry<Body>::body = &body;
    fiber_ctx.assign(block, ucontext_t{});
    fiber_state.assign(block, 0);
    std::vector<char *> stacks(block);
    for (unsigned i = 0; i < block; ++i) {
        stacks[i] = static_cast<char *>(mmap(nullptr, stack_size, PROT_READ | PROT_WRITE,
                                             MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0));
    }
#ifdef BBFFT_EMU_TSAN
    tsan_sched = __tsan_get_current_fiber();
    tsan_fiber.assign(block, nullptr);
    for (unsigned i = 0; i < block; ++i) {
        tsan_fiber[i] = __tsan_create_fiber(0);
    }
    fiber_epoch.assign(block, 0);
#endif
    gridDim_.x = grid;
    blockDim_.x = block;
    int rc = 0;
    for (unsigned b = 0; b < grid && rc == 0; ++b) {
        blockIdx_.x = b;
#ifdef BBFFT_EMU_TSAN
        for (unsigned i = 0; i < block; ++i) {
            __atomic_store_n(&fiber_epoch[i], 0, __ATOMIC_RELAXED);
        }
        __tsan_release(&group_token); // this group's set-up before its work-items
#endif
        for (unsigned i = 0; i < block; ++i) {
            getcontext(&fiber_ctx[i]);
            fiber_ctx[i].uc_stack.ss_sp = stacks[i];
            fiber_ctx[i].uc_stack.ss_size = stack_size;
            fiber_ctx[i].uc_link = &sched_ctx;
            makecontext(&fiber_ctx[i], &fiber_entry<Body>::run, 0);
            set_state(i, 0);
        }
        for (;;) {
            unsigned done = 0, at_barrier = 0;
            for (unsigned i = 0; i < block; ++i) {
                if (state(i) == 0) {
                    __atomic_store_n(&current, i, __ATOMIC_RELAXED);
#ifdef BBFFT_EMU_TSAN
                    if (__atomic_load_n(&fiber_epoch[i], __ATOMIC_RELAXED) >= epoch_token.size()) {
                        std::fprintf(stderr, "emulator: more barriers per work-group than "
                                             "race-check tokens\n");
                        std::abort();
                    }
                    __tsan_switch_to_fiber(tsan_fiber[i], __tsan_switch_to_fiber_no_sync);
#endif
                    swapcontext(&sched_ctx, &fiber_ctx[i]);
                }
            }
            for (unsigned i = 0; i < block; ++i) {
                done += state(i) == 2;
                at_barrier += state(i) == 1;
            }
            if (done == block) {
                break;
            }
            if (done != 0 && at_barrier != 0) {
                rc = 1; // some work-items finished while others wait at a barrier
                break;
            }
            ++barrier_count;
            for (unsigned i = 0; i < block; ++i) {
                set_state(i, 0);
            }
        }
#ifdef BBFFT_EMU_TSAN
        __tsan_acquire(&group_token); // the group's work before the next group's set-up
#endif
    }
#ifdef BBFFT_EMU_TSAN
    for (unsigned i = 0; i < block; ++i) {
        __tsan_destroy_fiber(tsan_fiber[i]);
    }
#endif
    for (unsigned i = 0; i < block; ++i) {
        munmap(stacks[i], stack_size);
    }
    return rc;
}

} // namespace emu

#if defined(__clang__)
namespace emu {
inline float mfma_a[16][64][8];
inline float mfma_b[16][64][8];

/// One v_mfma_f32_32x32x16_{f16,bf16} of a wave: A row / B column = lane % 32, inner index
/// 8 (lane / 32) + item; D item r of lane l = row 8 (r / 4) + 4 (l / 32) + r % 4, column
/// l % 32. Every lane of the work-group must call it (two work-group barriers).
template <typename V8, typename F16> F16 mfma_32x32x16(V8 a, V8 b, F16 c) {
    unsigned const lid = cur(), w = lid / 64u, l = lid % 64u;
    for (int e = 0; e < 8; ++e) {
        mfma_a[w][l][e] = static_cast<float>(a[e]);
        mfma_b[w][l][e] = static_cast<float>(b[e]);
    }
    __syncthreads();
    F16 d = c;
    unsigned const j = l % 32u;
    for (unsigned r = 0; r < 16; ++r) {
        unsigned const i = 8u * (r / 4u) + 4u * (l / 32u) + r % 4u;
        float s = d[r];
        for (unsigned k = 0; k < 16; ++k) {
            s += mfma_a[w][i + 32u * (k / 8u)][k % 8u] * mfma_b[w][j + 32u * (k / 8u)][k % 8u];
        }
        d[r] = s;
    }
    __syncthreads();
    return d;
}
} // namespace emu
#define __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, x, y, z) emu::mfma_32x32x16(a, b, c)
#define __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, x, y, z) emu::mfma_32x32x16(a, b, c)
#endif

#endif
