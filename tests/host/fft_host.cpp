// Host build of ipglasma_amd/csrc/fft_lds.hpp: the stages run one after the
// other over a whole row (all loads of a stage, then all stores, as the block
// barrier orders them on the device).
#include <cmath>
#include <vector>

#include "../../ipglasma_amd/csrc/fft_lds.hpp"

using namespace ipgfft;

template <int N, bool INV, int S>
static void host_stages(double2* x, const double2* tw) {
    if constexpr (S < nstages(N)) {
        constexpr int R = radix(N, S), NS = stride_ns(N, S), NB = N / R;
        std::vector<double2> v(N);
        for (int j = 0; j < NB; j++)
            for (int r = 0; r < R; r++) v[j * R + r] = x[j + r * NB];
        for (int j = 0; j < NB; j++) {
            double2 b[R];
            for (int r = 0; r < R; r++) b[r] = v[j * R + r];
            butterfly<N, R, NS, INV>(b, j, tw);
            const int o = out_index<N, R, NS>(j);
            for (int r = 0; r < R; r++) x[o + r * NS] = b[r];
        }
        host_stages<N, INV, S + 1>(x, tw);
    }
}

template <int N>
static int run(double* data, int inverse) {
    std::vector<double2> tw(N);
    for (int m = 0; m < N; m++) tw[m] = make_double2(std::cos(-2. * M_PI * m / N), std::sin(-2. * M_PI * m / N));
    double2* x = reinterpret_cast<double2*>(data);
    if (inverse)
        host_stages<N, true, 0>(x, tw.data());
    else
        host_stages<N, false, 0>(x, tw.data());
    return nstages(N);
}

extern "C" int ipgfft_host(int N, double* data, int inverse) {
    switch (N) {
        case 16: return run<16>(data, inverse);
        case 32: return run<32>(data, inverse);
        case 48: return run<48>(data, inverse);
        case 64: return run<64>(data, inverse);
        case 96: return run<96>(data, inverse);
        case 128: return run<128>(data, inverse);
        case 192: return run<192>(data, inverse);
        case 256: return run<256>(data, inverse);
        case 384: return run<384>(data, inverse);
        case 512: return run<512>(data, inverse);
        case 768: return run<768>(data, inverse);
        case 1024: return run<1024>(data, inverse);
    }
    return -1;
}
