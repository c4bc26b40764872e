// TEST INFRASTRUCTURE ONLY: runs one generated class-kernel body on the CPU
// (tests/emu/include/hip/hip_runtime.h) for tests/test_kernel_emulation.py.
// argv: input file, output file, team size T, teams to run, routed (0/1),
// optional FP32 error-check tolerance (the redo queue is then written after
// the output: its count, then the queued quartet indices, as ints).
// Input: nq, npair, npp, nout, then the IntcEriArgs arrays (quartets, pairs,
// primitive pairs, queue); the wave's teams run one after the other on one
// exactly sized LDS region and exactly sized device arrays, so AddressSanitizer
// reports any access outside them.  Output: the out array.
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
#include "intc_common.h"
#include "boys_tables.inc"
thread_local emu_dim3 threadIdx;
emu_dim3 blockIdx{0, 0, 0}, blockDim{64, 1, 1}, gridDim{1, 1, 1};
EmuTeam g_team;
double lds[1];
#define hipLaunchKernelGGL(...) ((void)0)
static inline hipError_t hipGetLastError() { return 0; }
#include CLASS_FILE
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
template <class T> T *rd(FILE *f, long n) { T *p = new T[n > 0 ? n : 1]; if (n && fread(p, sizeof(T), n, f) != (size_t)n) abort(); return p; }
int main(int argc, char **argv)
{
    FILE *f = fopen(argv[1], "rb");
    int hdr[4];
    if (fread(hdr, 4, 4, f) != 4) abort();
    const int nq = hdr[0], npair = hdr[1], npp = hdr[2], nout = hdr[3];
    const int T = atoi(argv[3]), nteams = atoi(argv[4]), routed = atoi(argv[5]);
    IntcEriArgs a;
    memset(&a, 0, sizeof(a));
    a.q_bra = rd<int>(f, nq);
    a.q_ket = rd<int>(f, nq);
    a.q_off = rd<long long>(f, nq);
    a.q_str = rd<unsigned long long>(f, nq);
    a.nq = nq;
    a.team_groups = T;
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
    int *nqdev = new int[1];
    nqdev[0] = nq;
    if (routed) { a.q_idx = qidx; a.nq_dev = nqdev; }
    fclose(f);
    double *cheb = new double[sizeof(intc_boys_cheb_host) / 8];
    memcpy(cheb, intc_boys_cheb_host, sizeof(intc_boys_cheb_host));
    a.boys_cheb = cheb;
    double *out = new double[nout];
    for (int i = 0; i < nout; ++i) out[i] = 12345.0;
    a.out = out;
    a.sph_builtin = 1;
    int *redo_q = new int[nq > 0 ? nq : 1];
    int *redo_n = new int[1];
    redo_n[0] = 0;
    if (argc > 6) {
        a.fp32_tol = atof(argv[6]);
        a.redo_q = redo_q;
        a.redo_n = redo_n;
    }
    const int G = CAT(KNAME, _G);
    const int TL = G * T, TPW = 64 / TL;
    const int reg = CAT(KNAME, _reg)(T);
    double *wlds = new double[reg];   // one wave's region (exact size: ASan sees overruns)
    int slot[64];
    g_team.slot = slot;
    // the wave's teams run one after the other (they share no LDS while the
    // wave is in lockstep); each team's lanes are threads meeting at barriers
    const int qstep = TPW;
    for (int tw = 0; tw < std::min(nteams, TPW); ++tw) {
        std::barrier<> bar(TL);
        g_team.bar = &bar;
        std::vector<std::thread> th;
        for (int l = tw * TL; l < (tw + 1) * TL; ++l)
            th.emplace_back([&, l]() {
                threadIdx.x = l;
                CAT(KNAME, _body)<0, 1>(a, wlds, 0, qstep, nq);
            });
        for (auto &t : th) t.join();
    }
    f = fopen(argv[2], "wb");
    fwrite(out, 8, nout, f);
    if (argc > 6) {
        fwrite(redo_n, 4, 1, f);
        fwrite(redo_q, 4, redo_n[0], f);
    }
    fclose(f);
    printf("emulated %d quartets, G=%d, T=%d, teams %d, routed %d, reg=%d doubles\n", nq, G, T, std::min(nteams, TPW), routed, reg);
    return 0;
}
