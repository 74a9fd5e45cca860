// Host build of ipglasma_amd/csrc/fft_lds.hpp: the stages run one after the
// other over whole rows held in the device's LDS layout (all loads of a
// stage, then all stores, as the block barrier orders them on the device).
#include <cmath>
#include <set>
#include <vector>

#include "../../ipglasma_amd/csrc/fft_lds.hpp"

using namespace ipgfft;

template <int N, bool INV, int S>
static void host_stages(double2* x, int nrow, const double2* tw) {
    if constexpr (S < nstages(N)) {
        constexpr int R = radix(N, S), NS = stride_ns(N, S), NB = N / R, RS = row_stride(N);
        for (int row = 0; row < nrow; row++) {
            std::vector<double2> v(N);
            for (int j = 0; j < NB; j++)
                for (int r = 0; r < R; r++) v[j * R + r] = x[row * RS + pad(j + r * NB)];
            for (int j = 0; j < NB; j++) {
                double2 b[R];
                for (int r = 0; r < R; r++) b[r] = v[j * R + r];
                butterfly<N, R, NS, INV>(b, j, tw);
                const int o = out_index<N, R, NS>(j);
                for (int r = 0; r < R; r++) x[row * RS + pad(o + r * NS)] = b[r];
            }
        }
        host_stages<N, INV, S + 1>(x, nrow, tw);
    }
}

// every (row, i) of nrow rows has its own slot inside lds_slots(N, nrow)
template <int N>
static bool layout_ok(int nrow) {
    std::set<int> used;
    for (int row = 0; row < nrow; row++)
        for (int i = 0; i < N; i++) {
            const int a = at(N, row, i);
            if (a < 0 || a >= lds_slots(N, nrow) || !used.insert(a).second) return false;
        }
    return true;
}

// data: nrow rows of N complex values, transformed in place
template <int N>
static int run(double* data, int nrow, int inverse) {
    if (!layout_ok<N>(4) || !layout_ok<N>(cols_per_block(N)) || !layout_ok<N>(nrow)) return -2;
    std::vector<double2> tw(N);
    for (int m = 0; m < N; m++) tw[m] = make_double2(std::cos(-2. * M_PI * m / N), std::sin(-2. * M_PI * m / N));
    std::vector<double2> lds(lds_slots(N, nrow), make_double2(NAN, NAN));
    double2* x = reinterpret_cast<double2*>(data);
    for (int row = 0; row < nrow; row++)
        for (int i = 0; i < N; i++) lds[at(N, row, i)] = x[row * N + i];
    if (inverse)
        host_stages<N, true, 0>(lds.data(), nrow, tw.data());
    else
        host_stages<N, false, 0>(lds.data(), nrow, tw.data());
    for (int row = 0; row < nrow; row++)
        for (int i = 0; i < N; i++) x[row * N + i] = lds[at(N, row, i)];
    return nstages(N);
}

extern "C" int ipgfft_host(int N, double* data, int nrow, int inverse) {
    switch (N) {
        case 16: return run<16>(data, nrow, inverse);
        case 32: return run<32>(data, nrow, inverse);
        case 48: return run<48>(data, nrow, inverse);
        case 64: return run<64>(data, nrow, inverse);
        case 96: return run<96>(data, nrow, inverse);
        case 128: return run<128>(data, nrow, inverse);
        case 192: return run<192>(data, nrow, inverse);
        case 256: return run<256>(data, nrow, inverse);
        case 384: return run<384>(data, nrow, inverse);
        case 512: return run<512>(data, nrow, inverse);
        case 768: return run<768>(data, nrow, inverse);
        case 1024: return run<1024>(data, nrow, inverse);
    }
    return -1;
}
