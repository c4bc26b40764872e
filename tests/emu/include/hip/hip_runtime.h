// TEST INFRASTRUCTURE ONLY (tests/test_kernel_emulation.py): a host stand-in
// for the HIP device subset that the generated class-kernel bodies use, so a
// body can be compiled with g++ -fsanitize=address and run on the CPU.  One
// OS thread per lane of one team; every cross-lane read (ds_bpermute) and
// wave barrier is a barrier of the team's lanes, which keeps them in
// lockstep wherever the kernel exchanges data between lanes.  The product
// never includes this file.
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
inline int __double2loint(double d) { uint64_t u; memcpy(&u, &d, 8); return (int)(uint32_t)u; }
inline int __double2hiint(double d) { uint64_t u; memcpy(&u, &d, 8); return (int)(uint32_t)(u >> 32); }
inline double __hiloint2double(int hi, int lo) { uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo; double d; memcpy(&d, &u, 8); return d; }
inline float __int_as_float(int i) { float f; memcpy(&f, &i, 4); return f; }
inline int __float_as_int(float f) { int i; memcpy(&i, &f, 4); return i; }
using std::fma; using std::sqrt; using std::exp; using std::round; using std::fmax; using std::fabs;
