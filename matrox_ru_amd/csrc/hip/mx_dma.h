// Direct global -> LDS loads (gfx950 LDS-DMA) and the waits and barriers
// of the pipelines built on them.
//
// global_load_lds_dword / _dwordx4 write lane l's 4 or 16 bytes to
// LDS[M0 + 4l or 16l] without passing through VGPRs; completion is counted
// by vmcnt like any vector load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mx {

typedef __attribute__((address_space(3))) void lds_void_t;
typedef const __attribute__((address_space(1))) void glb_void_t;

// LDS byte address of a __shared__ pointer
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) char *)p;
}

// s_waitcnt immediate for vmcnt(v) with expcnt and lgkmcnt left alone
constexpr int vmcnt_enc(int v) { return 0xF70 | (v & 15) | ((v >> 4) << 14); }

// wait until at most j * PER of this wave's LDS-DMA loads are outstanding
// (j <= J, the chunks that may stay in flight; vmcnt takes an immediate, so
// each j has its own instruction)
template <int PER, int J>
__device__ __forceinline__ void wait_dma(int j) {
    static_assert(J * PER < 64, "vmcnt is 6 bits");
    if constexpr (J == 0) {
        __builtin_amdgcn_s_waitcnt(vmcnt_enc(0));
    } else {
        if (j >= J) __builtin_amdgcn_s_waitcnt(vmcnt_enc(J * PER));
        else wait_dma<PER, J - 1>(j);
    }
}

// workgroup barrier without the memory fences of __syncthreads (which would
// wait for every outstanding prefetch); the signal fences keep the compiler
// from moving LDS accesses across it
__device__ __forceinline__ void wg_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

}  // namespace mx
