// ipg_internal.hpp -- host-side launchers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/ipg.h"
#include "ipg_device.hpp"

// sets the message ipg_last_error returns on the calling thread
void ipg_record_error(const std::string& msg);

// doubles of end-of-evolution observables per event (EvDesc::obs)
constexpr int kObsDoubles = 32;
// EvDesc::obs (the k_ecc_mid / k_ecc_finish layout) -> the Landau-frame and
// eccentricity fields of ipg_observables
inline void ipg_obs_from_device(const double* o, ipg_observables* q) {
    for (int n = 0; n < 6; n++) {
        q->eccentricity[n] = o[n];
        q->Psi[n] = o[6 + n];
    }
    q->area = o[12];
    q->Rbar = o[13];
    q->avg_eps = o[14];
    q->avg_qs2A_qs2B = o[15];
    q->rms_r = o[16];
    q->centroid[0] = o[17];
    q->centroid[1] = o[18];
    for (int k = 0; k < 3; k++) q->u_rms[k] = std::sqrt(o[19 + k] / o[22]);
    q->landau_failures = (int)o[23];
    q->landau_no_timelike = (int)o[24];
}

// site-parallel helpers: nblk = total blocks over all events of the launch
void ipg_launch_step0(const EvDesc* d_evs, int nev, int nblk, hipStream_t st);
// variant 0: register budget for 1 wave per SIMD, 1: for 2, 2: 4 x 64 site
// tiles along y, 3: the tiles down 64-wide column strips (same results)
void ipg_launch_step(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st, int variant = 0);
// the step for the blocks [block0, block0 + nblocks) of one event only
void ipg_launch_step_event(const EvDesc* d_evs, int nev, int block0, int nblocks, int it, hipStream_t st,
                           int variant = 0);
void ipg_launch_tmunu(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st);
void ipg_launch_gauss(const EvDesc* d_evs, int nev, int nblk, int it, hipStream_t st);
void ipg_launch_finish(const EvDesc* d_evs, int nev, int it, double* d_out, hipStream_t st);
// Landau frame + eccentricities of the events that end at it (all events if
// all != 0)
void ipg_launch_landau(const EvDesc* d_evs, int nev, int nblk, int it, int all, hipStream_t st);

// colour charge
void ipg_launch_tp(const EvDesc* d_evs, int nev, int nblk, hipStream_t st);
void ipg_launch_qs(const EvDesc* d_evs, int nev, int nblk, const double* d_Tlist, const double* d_Qs2,
                   int* d_flag_host_visible, hipStream_t st);
void ipg_launch_stats(const EvDesc* d_evs, int nev, int nblk, double* d_out, hipStream_t st);

// Wilson lines
struct FftPlanDev {
    int N;
    int logN;
    int C;                 // columns per block in the column pass
    const double2* tw;     // exp(-2 pi i m / N), m < N/2
    const double* s2;      // 4 sin^2(pi q / N), q < N
};
// launch shapes of the three setV passes (0: 256-thread blocks, 1: 512; for
// the forward row and the column passes 2: one wavefront per transformed
// row, synchronising itself only; for the inverse row pass 2 and 3: the
// next slice prefetched into registers, in 512-thread blocks and in
// 256-thread blocks of wavefront-synchronised transforms; inverse row 4, 5
// and forward row 3: shapes 0, 1 and 2 with the twiddles read from LDS; the
// same arithmetic) and the event pairs that bracket each when timed
struct SetvVariant {
    int fwd = 0, cols = 0, inv = 0;
};
struct SetvEvents {
    hipEvent_t fwd[2] = {nullptr, nullptr};
    hipEvent_t cols[2] = {nullptr, nullptr};
    hipEvent_t inv[2] = {nullptr, nullptr};
};
// slices [k0, k0 + kc) of every event of the launch (one lattice size)
void ipg_launch_setv_chunk(const EvDesc* d_evs, int nev, const FftPlanDev& plan, int k0, int kc,
                           const SetvVariant& v, const SetvEvents& t, hipStream_t st);
void ipg_launch_unit_v(const EvDesc* d_evs, int nev, int nblk, hipStream_t st);

// tau = 0 fields
void ipg_launch_links(const EvDesc* d_evs, int nev, int nblk, hipStream_t st);
void ipg_launch_newton(const EvDesc* d_evs, int nev, int nlinkblk, int* d_counter, int variant, hipStream_t st);
void ipg_launch_newton_stats(const EvDesc* d_evs, int nev, int nlinkblk, double* d_out, hipStream_t st);
void ipg_launch_pi_init(const EvDesc* d_evs, int nev, int nblk, hipStream_t st);
void ipg_launch_zero_fields(const EvDesc* d_evs, int nev, int nblk, hipStream_t st);
void ipg_launch_rho_normals(const EvDesc* d_evs, int nucleus, int k, double* d_out, hipStream_t st);

// reverse mode of dE/deta through the leapfrog (one event, evs[0]); the
// step's cotangents travel in pending sets of IPG_PEND_PLANES planes (own
// cotangents plus what each site sends to its neighbours), the loss fills
// one (zeroed by the caller), the kick turns the last into a compact set of
// 108 planes
#define IPG_PEND_PLANES 252
void ipg_adj_state_save(const EvDesc* d_evs, long long S, int parity, double* slot, hipStream_t st);
void ipg_adj_state_load(const EvDesc* d_evs, long long S, int parity, const double* slot, hipStream_t st);
void ipg_adj_zero(double* p, long long n, hipStream_t st);
void ipg_adj_loss(const EvDesc* d_evs, long long S, const double* xs, int it, double yb, double* cout,
                  hipStream_t st);
// variant 0: three kernels (own-field planes, link outboxes, phi outboxes),
// 1: one kernel running the three passes per site with the partial sums of
// the link cotangents in LDS, 2: the first kernel, then passes 2 and 3 as
// one with those sums in LDS (the same results)
void ipg_adj_step(const EvDesc* d_evs, long long S, const double* xb, const double* xa, int it, const double* cin,
                  double* cout, hipStream_t st, int variant = 0);
// T^{mu nu} at it in reverse: tb[10][S] cotangents of the stored components,
// field cotangents of the state xs into a zeroed pending set
void ipg_adj_tmunu(const EvDesc* d_evs, long long S, const double* xs, int it, const double* tb, double* pout,
                   hipStream_t st);
// d eps_2 / dT at it (after ipg_launch_landau of the same it) into tb[10][S]
void ipg_adj_ecc2(const EvDesc* d_evs, long long S, int it, double* tb, hipStream_t st);
// tb[k][s] = w[k] (device array of 10 weights)
void ipg_adj_fill_tb(long long S, const double* w, double* tb, hipStream_t st);
void ipg_adj_kick(const EvDesc* d_evs, long long S, const double* x0, const double* cin, double* cout,
                  hipStream_t st);
void ipg_adj_pi_init(const EvDesc* d_evs, long long S, const double* s0, const double* vab, const double* cs0,
                     double* ulb, double* db, hipStream_t st);
void ipg_adj_links_newton(const EvDesc* d_evs, long long S, const double* s0, const double* vab, const double* ulb,
                          const double* db, double* vb, hipStream_t st);
// setV in reverse for slices [k0, k0+kc) of event 0: vv/vbar are the current
// Wilson lines and their cotangents (2 nuclei x 18 planes), updated in place;
// grad[2][Ny][S][8] receives d/d rho; v: launch shapes of the recomputed
// forward row pass and of the column passes (as ipg_launch_setv_chunk);
// recompute false: the chunk's transform buffer already holds the filtered
// fields of the forward pass
void ipg_adj_setv_chunk(const EvDesc* d_evs, const FftPlanDev& plan, int k0, int kc, double* vv, double* vbar,
                        double* grad, const SetvVariant& v, hipStream_t st, bool recompute = true);


// Coulomb gauge fixing and gluon spectra (one event, evs[0]; par: the buffer
// holding the current U and phi).  A step computes chi = Im tr(div A t^a)
// and its row transform into buf (4 complex planes) with the residual's row
// partials in part[N], applies the stopping rule to the state gs[4]
// (stopped, step index, residual of the last tenth step, last residual),
// then filters, transforms back into the site rotations G (18 planes) and
// applies them to the fields; once gs says stopped every kernel returns at
// once.  ipg_launch_gf_reset starts gs.
void ipg_launch_gf_reset(double* gs, hipStream_t st);
void ipg_launch_gf_step(const EvDesc* d, int par, int N, const FftPlanDev& pl, double2* buf, double* G,
                        double* part, double* gs, hipStream_t st);
// the 2D transform of the Hermitian field F (9 planes) into buf (5 complex
// planes) and the per-block sums of n(k) = 2/(omega N^2) pref Re tr F(k)F(-k),
// omega n(k) ahbarc and the k_T histogram into part (ipg_mult_slots() doubles
// per block of 256 momenta)
void ipg_launch_mult(const EvDesc* d, int N, const FftPlanDev& pl, const double* F, double2* buf, double pref,
                     double ahbarc, double* part, hipStream_t st);
int ipg_mult_slots();
