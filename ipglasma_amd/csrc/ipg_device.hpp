// ipg_device.hpp -- per-event device descriptors shared by all kernels.
//
// One descriptor per event lives in device memory.  Every kernel launch
// covers a list of events: event e owns blocks [block0, block0 + nblocks) of
// the launch, so the plain path (one event per launch sequence) and the
// batched path (all events in one launch sequence) run the same kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su3.hpp"

#define IPG_BLOCK 256

struct EvDesc {
    // lattice
    int N;
    int Ny;
    long long S;
    int itmax;
    int block0;   // first block of this event in site-parallel launches
    int nblocks;  // ceil(S / IPG_BLOCK) padded to a multiple of 8
    int lblock0;  // first block in link-parallel launches (2S links)
    int lnblocks;  // ceil(2S / IPG_BLOCK) padded to a multiple of 8
    int active;   // 0: skipped by every launch (failed or finished)
    // physics (lattice units unless noted)
    double g, L, a, dtau, m_lat, uvdamp_lat;
    // fields: SoA, 18 planes of S doubles each
    double* U[2][2];  // [parity][0:x, 1:y]
    double* phi[2];   // [parity]
    double* E1;
    double* E2;
    double* pi;
    // where k_step writes the updated E1, E2, pi: the same planes (in place),
    // except for the reverse mode's recomputed steps, which go from one
    // stored state to the next
    double* E1o;
    double* E2o;
    double* pio;
    // Wilson lines V_A, V_B (alias the parity-1 link buffers during init)
    double* VA;
    double* VB;
    // links U^(1)_x, U^(1)_y, U^(2)_x, U^(2)_y during init (alias E1, E2,
    // phi[0], phi[1])
    double* L1x;
    double* L1y;
    double* L2x;
    double* L2y;
    // colour charge
    double* TpA;
    double* TpB;
    double* g2A;
    double* g2B;
    double* yB;      // final rapidity of the B fixed point per site
    int* check2;     // site inside both nuclei (Init.cpp:1255-1301)
    // Glauber nucleons (x, y, phi, collided) for A then B
    double* nuc;
    int A1, A2;
    // Q_s / g2mu2 settings
    double BG, rmax, QsmuRatio, QsmuRatioB, xFactor, roots, rapidity, xi;
    double rcut;  // sqrt(0.1 sigma_NN / pi) [fm] (Init.cpp:1464)
    int useFluctuatingx;
    int qs_flags;  // bit0: serial fallback needed
    // setV
    double2* fftbuf;     // [2 nuclei][kc slices][4 packed][S] complex
    int kc;              // slices per setV chunk
    const double* znorm;  // optional host-provided normals [2][Ny][S][8]
    uint64_t rho_seed;
    uint64_t newton_seed;
    int newton_polish;  // Newton iterations after the stopping rule holds
    // outputs
    double* T;        // 10 planes of S doubles
    double* partial;  // per-block reduction partials (32 slots of nblocks)
    int* newton_stat;  // per-link iteration counts (2S) and flags
    // Landau frame: eps, u^tau, u^x, u^y, u^eta, pi^{mu nu} (T's order)
    double* land;       // 15 planes of S doubles
    double* land_epsv;  // running eps of the reference after each site
    int* land_flag;     // bit0 eps set, bit1 u^mu set, bit2 QR failed
    double* obs;        // 32 doubles of end-of-evolution observables
};

// block -> (event, local block) for site-parallel launches
__device__ __forceinline__ int find_event(const EvDesc* __restrict__ ev, int nev, int b) {
    int e = 0;
    for (int k = 1; k < nev; k++)
        if (b >= ev[k].block0) e = k;
    return e;
}
__device__ __forceinline__ int find_event_links(const EvDesc* __restrict__ ev, int nev, int b) {
    int e = 0;
    for (int k = 1; k < nev; k++)
        if (b >= ev[k].lblock0) e = k;
    return e;
}

// Spread consecutive logical blocks over one XCD: hardware dispatches block b
// to XCD b % 8, so with n a multiple of 8 the bijection
// l = (b % 8) * (n / 8) + b / 8 keeps neighbouring lattice rows inside one
// XCD's L2.  Every event's block count is padded to a multiple of 8 and its
// block0 is a multiple of 8, so local and global block indices agree mod 8.
__device__ __forceinline__ int xcd_remap(int b, int n8) { return (b % 8) * (n8 / 8) + b / 8; }

__device__ __forceinline__ int clampi(int v, int N) { return v < 0 ? 0 : (v > N - 1 ? N - 1 : v); }
