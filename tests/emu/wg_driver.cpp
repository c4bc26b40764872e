// TEST INFRASTRUCTURE ONLY: runs one workgroup-kernel instantiation
// intc_eri_wg_k<VT, BM, LA, LB, LC, LD> on the CPU
// (tests/emu/include/hip/hip_runtime.h) for tests/test_wg_emulation.py.
// Compile-time: WG_FILE (the kernel source up to its launch code, with the
// dynamic LDS array declared as a pointer), EMU_VT, EMU_LA .. EMU_LD.
// argv: input file, output file, spherical (0/1), workgroups, routed (0/1),
// Boys mode (0/1).  Input: nq, npair, npp, nout, the IntcEriArgs arrays (as
// lane_driver.cpp reads them), the queue, then the S_l blob.  The workgroups
// run one after the other, each as 256 threads on one LDS region of exactly
// the size the host layout gives (filled with NaN bytes first, so a read of an
// entry no thread wrote shows in the output) and on exactly sized device
// arrays, so AddressSanitizer reports any access outside them.
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
#include <hip/hip_runtime.h>
thread_local emu_dim3 threadIdx;
emu_dim3 blockIdx{0, 0, 0}, blockDim{256, 1, 1}, gridDim{1, 1, 1};
EmuTeam g_team;
unsigned char *emu_wg_lds = nullptr;
constexpr hipError_t hipSuccess = 0, hipErrorInvalidValue = 1;
#define hipLaunchKernelGGL(...) ((void)0)
static inline hipError_t hipGetLastError() { return 0; }
#include "intc_common.h"
#include "boys_tables.inc"
#include WG_FILE
template <class T> T *rd(FILE *f, long n) { T *p = new T[n > 0 ? n : 1]; if (n && fread(p, sizeof(T), n, f) != (size_t)n) abort(); return p; }
int main(int argc, char **argv)
{
    FILE *f = fopen(argv[1], "rb");
    int hdr[4];
    if (fread(hdr, 4, 4, f) != 4) abort();
    const int nq = hdr[0], npair = hdr[1], npp = hdr[2], nout = hdr[3];
    const int sph = atoi(argv[3]), nblk = atoi(argv[4]), routed = atoi(argv[5]), bm = atoi(argv[6]);
    IntcEriArgs a;
    memset(&a, 0, sizeof(a));
    a.q_bra = rd<int>(f, nq);
    a.q_ket = rd<int>(f, nq);
    a.q_off = rd<long long>(f, nq);
    a.q_str = rd<unsigned long long>(f, nq);
    a.nq = nq;
    a.pair_off = rd<int>(f, npair);
    a.pair_np = rd<int>(f, npair);
    a.pair_nk = rd<int>(f, npair);
    a.pair_A = rd<double>(f, 3 * npair);
    a.pair_AB = rd<double>(f, 3 * npair);
    a.pp_zeta = rd<double>(f, npp);
    a.pp_o2z = rd<double>(f, npp);
    a.pp_px = rd<double>(f, npp);
    a.pp_py = rd<double>(f, npp);
    a.pp_pz = rd<double>(f, npp);
    a.pp_kappa = rd<double>(f, npp);
    int *qidx = rd<int>(f, nq);
    a.sph = rd<double>(f, intc_sph_off(8));
    fclose(f);
    int *nqdev = new int[1];
    nqdev[0] = nq;
    if (routed) { a.q_idx = qidx; a.nq_dev = nqdev; }
    double *cheb = new double[sizeof(intc_boys_cheb_host) / 8];
    memcpy(cheb, intc_boys_cheb_host, sizeof(intc_boys_cheb_host));
    a.boys_cheb = cheb;
    double *bsm = new double[sizeof(intc_boys_smallx_host) / 8], *bmd = new double[sizeof(intc_boys_mediumx_host) / 8];
    memcpy(bsm, intc_boys_smallx_host, sizeof(intc_boys_smallx_host));
    memcpy(bmd, intc_boys_mediumx_host, sizeof(intc_boys_mediumx_host));
    a.boys_small = bsm;
    a.boys_medium = bmd;
    a.boys_mode = bm;
    a.wg_cls[0] = EMU_LA; a.wg_cls[1] = EMU_LB; a.wg_cls[2] = EMU_LC; a.wg_cls[3] = EMU_LD; a.wg_cls[4] = sph;
    double *out = new double[nout];
    for (int i = 0; i < nout; ++i) out[i] = 12345.0;
    a.out = out;
    IntcWgCls k;
    if (wg2_layout<EMU_LA, EMU_LB, EMU_LC, EMU_LD>(sph, (int)sizeof(EMU_VT), 160 * 1024, &k)) { fprintf(stderr, "LDS over 160 KB\n"); return 3; }
    printf("lds_bytes %d\n", k.lds_bytes);
    emu_wg_lds = new unsigned char[k.lds_bytes];
    gridDim.x = nblk;
    for (int b = 0; b < nblk; ++b) {
        memset(emu_wg_lds, 0xff, k.lds_bytes);
        blockIdx.x = b;
        std::barrier<> bar(256);
        std::barrier<> wb[4] = {std::barrier<>(64), std::barrier<>(64), std::barrier<>(64), std::barrier<>(64)};
        g_block_bar = &bar;
        for (int w = 0; w < 4; ++w) g_wave_bar[w] = &wb[w];
        std::vector<std::thread> th;
        for (int t = 0; t < 256; ++t)
            th.emplace_back([&, t]() {
                threadIdx.x = t;
                if (bm)
                    intc_eri_wg_k<EMU_VT, 1, EMU_LA, EMU_LB, EMU_LC, EMU_LD>(a, k);
                else
                    intc_eri_wg_k<EMU_VT, 0, EMU_LA, EMU_LB, EMU_LC, EMU_LD>(a, k);
            });
        for (auto &t : th) t.join();
    }
    FILE *g = fopen(argv[2], "wb");
    fwrite(out, 8, nout, g);
    fclose(g);
    return 0;
}
