// nw_seq.h -- device helpers shared by the alignment kernels: 2-bit code access,
// the flag -> (s, e) rule and the reference's traceback decision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srm {

// ---------------------------------------------------------------------------
// sequence access
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t code_at(const uint32_t *__restrict__ words, int64_t idx)
{
    return (words[idx >> 4] >> ((uint32_t)(idx & 15) * 2u)) & 3u;
}

// 16 consecutive codes starting at code index idx (idx may be negative down to
// -SRM_CODE_PAD: the arrays carry padding)
__device__ __forceinline__ uint32_t fetch16(const uint32_t *__restrict__ words, int64_t idx)
{
    int64_t wi = idx >> 4;
    uint32_t sh = (uint32_t)(idx & 15) * 2u;
    uint32_t lo = words[wi];
    uint32_t hi = words[wi + 1];
    return __builtin_amdgcn_alignbit(hi, lo, sh);
}

// cuda-functions.cu:383-410: flags 2,3,6,7 free the start (s = k), flags 1,3,5,7
// free the end (e = k); every other value leaves both at 0
__device__ __forceinline__ int flag_s(int flag, int k) { return (flag <= 7 && (flag & 2)) ? k : 0; }
__device__ __forceinline__ int flag_e(int flag, int k) { return (flag <= 7 && (flag & 1)) ? k : 0; }

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// value-based traceback decision, cuda-functions.cu:121-276 (matrix-relative).
// returns 0 diag, 1 top, 2 left, 3 stop
__device__ __forceinline__ int decide(int ma, int mm, int v, int Dg, int Lf, int Tp, int wd, int ht,
                                      int L, int w, int s, int e, bool chars_equal)
{
    int corr = ((wd == L - 1 && ht >= w - 2 * e) || (s && wd == 0 && ht <= (s + 1) * 2)) ? mm : 0;
    if (Dg + mm == v) return 0;
    if (Dg + ma == v && chars_equal) return 0;
    if (Tp + mm - corr == v) return 1;
    if (Lf + mm == v) return 2;
    return 3;
}

} // namespace srm
