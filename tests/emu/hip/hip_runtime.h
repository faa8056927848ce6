// Host stand-in for the few HIP names used by saruman_amd/csrc/band_bits.hip, so
// that the kernel source can run on the CPU as test code: lanes run one after
// another (each kernel's lanes are independent; wave-uniform values are uniform
// anyway), launches and occupancy queries are no-ops.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__
#define amdgpu_waves_per_eu(...) unused

struct uint4 { uint32_t x, y, z, w; };
struct dim3 {
    uint32_t x, y, z;
    dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
typedef int hipError_t;
typedef void *hipStream_t;
#define hipSuccess 0
#define hipErrorInvalidValue 1
inline hipError_t hipGetLastError() { return 0; }
#define hipOccupancyMaxActiveBlocksPerMultiprocessor(b, f, t, s) (*(b) = 1, 0)
#define hipLaunchKernelGGL(...) ((void)0)

struct EmuIdx { uint32_t x, y, z; };
extern thread_local EmuIdx threadIdx, blockIdx, blockDim;

// gfx9 semantics: v_alignbit_b32 (64-bit funnel shift right, low 32 bits),
// v_bfe_i32 (signed bit-field extract), readfirstlane (uniform values here)
inline uint32_t __builtin_amdgcn_alignbit(uint32_t hi, uint32_t lo, uint32_t sh)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (sh & 31));
}
inline int __builtin_amdgcn_sbfe(int x, uint32_t off, uint32_t w)
{
    off &= 31;
    return (int)((int64_t)((uint64_t)(uint32_t)x << (64 - off - w)) >> (64 - w));
}
inline uint32_t __builtin_amdgcn_readfirstlane(uint32_t x) { return x; }
// v_bfe_u32 (unsigned bit-field extract: w bits from bit off, both taken mod 32)
inline uint32_t __builtin_amdgcn_ubfe(uint32_t x, uint32_t off, uint32_t w)
{
    off &= 31;
    w &= 31;
    if (w == 0) return 0;
    return (uint32_t)(((uint64_t)x >> off) & ((1ull << w) - 1ull));
}
// a wave vote, taken here over the one lane that runs: a kernel may use it only for
// choices that are correct for each lane that votes true
inline bool __all(bool x) { return x; }
inline int __builtin_amdgcn_readfirstlane(int x) { return x; }
