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
#define BBFFT_WAIT_ALL() ((void)0)
#define BBFFT_LDS_SYNC() __syncthreads()
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

namespace emu {
inline emu_dim3 threadIdx_, blockIdx_, gridDim_, blockDim_;
inline ucontext_t sched_ctx;
inline std::vector<ucontext_t> fiber_ctx;
inline std::vector<int> fiber_state; // 0 runnable, 1 at barrier, 2 done
inline unsigned current = 0;
inline std::size_t barrier_count = 0;
} // namespace emu

#define threadIdx (emu::threadIdx_)
#define blockIdx (emu::blockIdx_)
#define gridDim (emu::gridDim_)
#define blockDim (emu::blockDim_)

inline void __syncthreads() {
    unsigned me = emu::current;
    emu::fiber_state[me] = 1;
    swapcontext(&emu::fiber_ctx[me], &emu::sched_ctx);
}

namespace emu {

/// Runs body() for `block` work-items of each of `grid` work-groups. Returns 0 on success,
/// 1 if work-items of one group disagree on the number of barriers (divergent barrier).
template <typename Body> int launch(unsigned grid, unsigned block, Body body) {
    static std::size_t const stack_size = 512 * 1024;
    static Body *body_ptr;
    body_ptr = &body;
    fiber_ctx.assign(block, ucontext_t{});
    fiber_state.assign(block, 0);
    std::vector<char *> stacks(block);
    for (unsigned i = 0; i < block; ++i) {
        stacks[i] = static_cast<char *>(mmap(nullptr, stack_size, PROT_READ | PROT_WRITE,
                                             MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0));
    }
    gridDim_.x = grid;
    blockDim_.x = block;
    int rc = 0;
    for (unsigned b = 0; b < grid && rc == 0; ++b) {
        blockIdx_.x = b;
        for (unsigned i = 0; i < block; ++i) {
            getcontext(&fiber_ctx[i]);
            fiber_ctx[i].uc_stack.ss_sp = stacks[i];
            fiber_ctx[i].uc_stack.ss_size = stack_size;
            fiber_ctx[i].uc_link = &sched_ctx;
            makecontext(&fiber_ctx[i], (void (*)()) + [] {
                (*body_ptr)();
                fiber_state[current] = 2;
            }, 0);
            fiber_state[i] = 0;
        }
        for (;;) {
            unsigned done = 0, at_barrier = 0;
            for (unsigned i = 0; i < block; ++i) {
                if (fiber_state[i] == 0) {
                    current = i;
                    threadIdx_.x = i;
                    swapcontext(&sched_ctx, &fiber_ctx[i]);
                }
            }
            for (unsigned i = 0; i < block; ++i) {
                done += fiber_state[i] == 2;
                at_barrier += fiber_state[i] == 1;
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
                fiber_state[i] = 0;
            }
        }
    }
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
    unsigned const lid = threadIdx_.x, w = lid / 64u, l = lid % 64u;
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
