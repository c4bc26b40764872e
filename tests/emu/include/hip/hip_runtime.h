// TEST INFRASTRUCTURE ONLY (tests/test_kernel_emulation.py,
// tests/test_wg_emulation.py): a host stand-in for the HIP device subset that
// the generated class-kernel bodies and the workgroup kernels use, so they can
// be compiled with g++ -fsanitize=address and run on the CPU.  One OS thread
// per lane.  Class bodies: one team; every cross-lane read (ds_bpermute) and
// wave barrier is a barrier of the team's lanes, which keeps them in lockstep
// wherever the kernel exchanges data between lanes.  Workgroup kernels: one
// 256-thread block; __syncthreads is a barrier of the block and the FP64 MFMA
// builtin a barrier of the wave's 64 lanes around an exchange of the operands
// (the documented v_mfma_f64_16x16x4f64 lane layout).  The product never
// includes this file.
#pragma once
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __constant__
#define __shared__
#define __launch_bounds__(x)
#define __restrict__
typedef int hipError_t;
typedef void *hipStream_t;
struct emu_dim3 { int x, y, z; };
extern thread_local emu_dim3 threadIdx;
extern emu_dim3 blockIdx, blockDim, gridDim;
struct EmuTeam {
    std::barrier<> *bar;
    int *slot;
};
extern EmuTeam g_team;
inline int emu_bpermute(int addr, int v)
{
    g_team.slot[threadIdx.x] = v;
    g_team.bar->arrive_and_wait();
    int src = addr / 4;
    if (src < 0 || src >= 64) { fprintf(stderr, "bpermute source lane %d out of range\n", src); abort(); }
    int r = g_team.slot[src];
    g_team.bar->arrive_and_wait();
    return r;
}
#define __builtin_amdgcn_ds_bpermute(a, v) emu_bpermute((a), (v))
#define __builtin_amdgcn_wave_barrier() g_team.bar->arrive_and_wait()
#define __builtin_amdgcn_fence(a, b) ((void)0)
#define __builtin_amdgcn_sched_barrier(m) ((void)0)
// the hardware reciprocal square root is an estimate: stand in with a
// single-precision one, so the Newton steps that follow it are exercised
#define __builtin_amdgcn_rsq(x) ((double)(1.0f / sqrtf((float)(x))))
// device atomics: the lanes of the team are host threads
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int __double2loint(double d) { uint64_t u; memcpy(&u, &d, 8); return (int)(uint32_t)u; }
inline int __double2hiint(double d) { uint64_t u; memcpy(&u, &d, 8); return (int)(uint32_t)(u >> 32); }
inline double __hiloint2double(int hi, int lo) { uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo; double d; memcpy(&d, &u, 8); return d; }
inline float __int_as_float(int i) { float f; memcpy(&f, &i, 4); return f; }
inline int __float_as_int(float f) { int i; memcpy(&i, &f, 4); return i; }
// block barrier and the wave barriers of the MFMA exchange (set by the driver)
inline std::barrier<> *g_block_bar = nullptr;
inline std::barrier<> *g_wave_bar[4] = {nullptr, nullptr, nullptr, nullptr};
inline double g_mfma_a[256], g_mfma_b[256];
inline void __syncthreads() { g_block_bar->arrive_and_wait(); }
// the workgroup kernels' one vector type is four doubles (the MFMA accumulator)
#define ext_vector_type(n) vector_size((n) * sizeof(double))
typedef double emu_f64x4 __attribute__((vector_size(4 * sizeof(double))));
// D[i][j] += sum_k A[i][k] B[k][j] over the wave: lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16] and D[l / 16 + 4 v][l % 16], v = 0..3
inline emu_f64x4 emu_mfma_f64_16x16x4(double a, double b, emu_f64x4 c)
{
    const int t = threadIdx.x, l = t & 63, w = t >> 6, base = t - l;
    g_mfma_a[t] = a;
    g_mfma_b[t] = b;
    g_wave_bar[w]->arrive_and_wait();
    const int j = l % 16;
    emu_f64x4 d = c;
    for (int v = 0; v < 4; ++v) {
        const int i = l / 16 + 4 * v;
        for (int k = 0; k < 4; ++k) d[v] = std::fma(g_mfma_a[base + i + 16 * k], g_mfma_b[base + j + 16 * k], d[v]);
    }
    g_wave_bar[w]->arrive_and_wait();
    return d;
}
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) emu_mfma_f64_16x16x4((a), (b), (c))
inline int min(int a, int b) { return a < b ? a : b; }
struct uint2 { unsigned x, y; };
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
using std::fma; using std::sqrt; using std::exp; using std::round; using std::fmax; using std::fabs;
