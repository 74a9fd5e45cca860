// host.cpp -- input parsing, MT19937-64, nucleus table, nucleon sampling.
//
// Follows (draw order is part of the contract: nucleon positions and the
// binary-collision list are bit-exact targets):
//   Setup.cpp:24-104, main.cpp:448-578   input keys, truncation/rounding rules
//   main.cpp:531-534, Evolution.cpp:396  derived dtau and step count
//   Random.cpp:67-204                    MT19937-64 and its transforms
//   Glauber.cpp:19-287                   nucleus constants
//   Init.cpp:2320-2345                   impact parameter
//   Init.cpp:65-439, 2721-3178           sampleTA and its samplers
//   Init.cpp:868-901, 1117-1216          reaction plane, shift, wounding
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <numeric>

namespace ipg {

// ---- MT19937-64 ---------------------------------------------------------------
void MT64::seed(uint64_t s) {
    mt_[0] = s;
    for (int i = 1; i < 312; i++) mt_[i] = 6364136223846793005ULL * (mt_[i - 1] ^ (mt_[i - 1] >> 62)) + (uint64_t)i;
    mti_ = 312;
    iset_ = 0;
    gset_ = 0.;
}

uint64_t MT64::next() {
    constexpr uint64_t kA = 0xB5026F5AA96619E9ULL, kUp = 0xFFFFFFFF80000000ULL, kLo = 0x7FFFFFFFULL;
    if (mti_ >= 312) {
        if (mti_ == 313) seed(5489ULL);
        for (int i = 0; i < 312; i++) {
            uint64_t y = (mt_[i] & kUp) | (mt_[(i + 1) % 312] & kLo);
            mt_[i] = mt_[(i + 156) % 312] ^ (y >> 1) ^ ((y & 1ULL) ? kA : 0ULL);
        }
        mti_ = 0;
    }
    uint64_t x = mt_[mti_++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}
double MT64::real1() { return (double)(next() >> 11) * (1.0 / 9007199254740991.0); }
double MT64::real2() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
double MT64::real3() { return ((double)(next() >> 12) + 0.5) * (1.0 / 4503599627370496.0); }
double MT64::gauss() {
    if (iset_) {
        iset_ = 0;
        return gset_;
    }
    double v1, v2, r2;
    do {
        v1 = 2.0 * real3() - 1.0;
        v2 = 2.0 * real3() - 1.0;
        r2 = v1 * v1 + v2 * v2;
    } while (r2 > 1. || r2 == 0.);
    const double f = std::sqrt(-2.0 * std::log(r2) / r2);
    gset_ = v1 * f;
    iset_ = 1;
    return v2 * f;
}

// ---- nucleus table ---------------------------------------------------------------
namespace {
struct Row {
    const char* name;
    int A, Z;
    double R, a, b2, b3, b4;
    int user_b2;  // beta2 taken from the input (U, Xe)
};
// Glauber.cpp:19-245: Woods-Saxon nuclei, the proton and the deuteron
// (Hulthen a = 0.228 fm^-1 in the a column, R_WS = 1 its starting
// normalisation); He3 needs a sampler that is not on this path
const Row kNuclei[] = {
    {"Au", 197, 79, 6.37, 0.535, -0.13, 0., -0.03, 0}, {"Pb", 208, 82, 6.62, 0.546, 0., 0., 0., 0},
    {"p", 1, 1, 1., 1., 0., 0., 0., 0},                {"d", 2, 1, 1.0, 0.228, 0., 0., 0., 0},
    {"C", 12, 6, 2.44, 1.635, 0., 0., 0., 0},          {"O", 16, 8, 2.608, 0.513, -0.01, 0., -0.122, 0},
    {"Ne", 20, 10, 2.8, 0.57, 0., 0., 0., 0},           {"S", 32, 16, 2.54, 2.191, 0., 0., 0., 0},
    {"Ar", 40, 18, 3.61, 0.516, 0.1668, 0., 0.00695, 0}, {"W", 184, 74, 6.51, 0.535, 0., 0., 0., 0},
    {"Al", 27, 13, 3.07, 0.519, 0., 0., 0., 0},         {"Ca", 40, 20, 3.766, 0.586, 0., 0., 0., 0},
    {"Cu", 63, 29, 4.163, 0.606, 0.162, 0., 0.006, 0},  {"Fe", 56, 26, 4.106, 0.519, 0., 0., 0., 0},
    {"Pt", 195, 78, 6.78, 0.54, 0., 0., 0., 0},         {"U", 238, 92, 6.81, 0.55, 0., 0., 0.093, 1},
    {"Ru", 96, 44, 5.085, 0.46, 0.158, 0., 0., 0},      {"Zr", 96, 40, 5.02, 0.46, 0., 0., 0., 0},
    {"Xe", 129, 54, 5.42, 0.57, 0., 0., -0.003, 1},
};
}  // namespace

bool find_nucleus(const std::string& name, const ipg_params& p, NucleusData* out) {
    for (const Row& r : kNuclei) {
        if (name != r.name) continue;
        NucleusData d;
        d.A = r.A;
        d.Z = r.Z;
        d.R = r.R;
        d.a = r.a;
        d.beta2 = r.user_b2 ? p.beta2 : r.b2;
        d.beta3 = r.b3;
        d.beta4 = r.b4;
        d.gamma = 0.;
        d.dR = 0.;
        d.da = 0.;
        if (p.setWSDeformParams) {
            d.R = p.R_WS;
            d.a = p.a_WS;
            d.beta2 = p.beta2;
            d.beta3 = p.beta3;
            d.beta4 = p.beta4;
            d.gamma = p.gamma;
            d.dR = p.dR_np;
            d.da = p.da_np;
        }
        d.force_dmin = p.force_dmin_flag;
        d.d_min = p.d_min;
        d.w = r.A == 2 ? 1.18 : 0.;  // Glauber.cpp:68-77 (not overridden)
        *out = d;
        return true;
    }
    return false;
}

// ---- nucleon sampling ---------------------------------------------------------------
namespace {

double fermi(double r, double R, double a) { return 1. / (1. + std::exp((r - R) / a)); }

double ws_radius(MT64& rng, double a, double R) {
    const double rmax = R + 10. * a;
    double r;
    do {
        r = rmax * std::pow(rng.real3(), 1.0 / 3.0);
    } while (rng.real3() > fermi(r, R, a));
    return r;
}

double ylm(int l, double ct) {
    if (l == 2) return (3.0 * ct * ct - 1.0) * 0.31539156525252005;
    if (l == 3) return (5.0 * ct * ct * ct - 3.0 * ct) * 0.3731763325901154;
    return (35.0 * ct * ct * ct * ct - 30.0 * ct * ct + 3.0) * 0.10578554691520431;
}

void recentre(std::vector<double>& x, std::vector<double>& y, std::vector<double>& z) {
    double mx = 0., my = 0., mz = 0.;
    for (size_t i = 0; i < x.size(); i++) {
        mx += x[i];
        my += y[i];
        mz += z[i];
    }
    mx /= (double)x.size();
    my /= (double)y.size();
    mz /= (double)z.size();
    for (size_t i = 0; i < x.size(); i++) {
        x[i] -= mx;
        y[i] -= my;
        z[i] -= mz;
    }
}

void generate_nucleus(MT64& rng, const NucleusData& d, std::vector<Nucleon>& out) {
    const int A = d.A, Z = d.Z;
    std::vector<double> x(A, 0.), y(A, 0.), z(A, 0.);
    std::vector<int> proton(A, 1);
    const bool spherical = std::abs(d.beta2) < 1e-15 && std::abs(d.beta4) < 1e-15 && std::abs(d.beta3) < 1e-15 &&
                           std::abs(d.gamma) < 1e-15;
    if (spherical || (!d.force_dmin && std::abs(d.gamma) <= 1e-15)) {
        // radii (and polar angles) first, sorted by radius, then placed
        // with the d_min rejection (Init.cpp:2749-2816, 2833-2904)
        std::vector<double> r(A, 0.), ct(A, 0.);
        std::vector<int> idx(A);
        for (int i = 0; i < A; i++) {
            const double ai = i < Z ? d.a : d.a + d.da, Ri = i < Z ? d.R : d.R + d.dR;
            if (spherical) {
                r[i] = ws_radius(rng, ai, Ri);
            } else {
                const double rmax = Ri + 10. * ai;
                double Rth;
                do {
                    r[i] = rmax * std::pow(rng.real3(), 1.0 / 3.0);
                    ct[i] = 1.0 - 2.0 * rng.real3();
                    Rth = Ri * (1.0 + d.beta2 * ylm(2, ct[i]) + d.beta3 * ylm(3, ct[i]) + d.beta4 * ylm(4, ct[i]));
                } while (rng.real3() > fermi(r[i], Rth, ai));
            }
            idx[i] = i;
        }
        std::stable_sort(idx.begin(), idx.end(), [&r](int i1, int i2) { return r[i1] < r[i2]; });
        std::vector<double> rs(A);
        for (int i = 0; i < A; i++) rs[i] = r[idx[i]];
        const double dmin2 = d.d_min * d.d_min;
        for (int i = 0; i < A; i++) {
            const double ri = rs[i];
            const double th_fixed = spherical ? 0. : std::acos(ct[idx[i]]);
            double xi = 0, yi = 0, zi = 0;
            int iter = 0, reject;
            do {
                iter++;
                reject = 0;
                const double phi = 2. * M_PI * rng.real3();
                const double th = spherical ? std::acos(1. - 2. * rng.real3()) : th_fixed;
                xi = ri * std::sin(th) * std::cos(phi);
                yi = ri * std::sin(th) * std::sin(phi);
                zi = ri * std::cos(th);
                for (int j = i - 1; j >= 0; j--) {
                    if ((ri - rs[j]) * (ri - rs[j]) > dmin2) break;
                    const double dsq = (xi - x[j]) * (xi - x[j]) + (yi - y[j]) * (yi - y[j]) + (zi - z[j]) * (zi - z[j]);
                    if (dsq < dmin2) {
                        reject = 1;
                        break;
                    }
                }
            } while (reject == 1 && iter < 100);
            x[i] = xi;
            y[i] = yi;
            z[i] = zi;
            // the spherical sampler marks every nucleon a proton (Init.cpp:2809-2813)
            proton[i] = spherical ? 1 : (idx[i] < Z);
        }
    } else {
        // Init.cpp:2906-2982 (forced d_min) and 2984-3037 (triaxial)
        const double rmax = d.R + d.dR + 10. * (d.a + d.da);
        const double dmin2 = d.d_min * d.d_min;
        for (int i = 0; i < A; i++) {
            const double Ri = i >= Z ? d.R + d.dR : d.R, ai = i >= Z ? d.a + d.da : d.a;
            bool again;
            double xi, yi, zi;
            do {
                double r, ct, phi, Rth;
                do {
                    r = rmax * std::pow(rng.real3(), 1.0 / 3.0);
                    ct = 1.0 - 2.0 * rng.real3();
                    phi = 2. * M_PI * rng.real3();
                    const double y22 = (1.0 - ct * ct) * std::cos(2. * phi) * 0.5462742152960397;
                    Rth = Ri * (1.0 + d.beta2 * (std::cos(d.gamma) * ylm(2, ct) + std::sin(d.gamma) * y22) +
                                d.beta3 * ylm(3, ct) + d.beta4 * ylm(4, ct));
                } while (rng.real3() > fermi(r, Rth, ai));
                const double st = std::sqrt(1. - ct * ct);
                xi = r * st * std::cos(phi);
                yi = r * st * std::sin(phi);
                zi = r * ct;
                again = false;
                if (d.force_dmin) {
                    for (int j = i - 1; j >= 0; j--) {
                        const double dsq = (xi - x[j]) * (xi - x[j]) + (yi - y[j]) * (yi - y[j]) + (zi - z[j]) * (zi - z[j]);
                        if (dsq < dmin2) {
                            again = true;
                            break;
                        }
                    }
                }
            } while (again);
            x[i] = xi;
            y[i] = yi;
            z[i] = zi;
            proton[i] = i < Z;
        }
    }
    recentre(x, y, z);
    out.resize(A);
    for (int i = 0; i < A; i++) out[i] = Nucleon{x[i], y[i], z[i], std::atan2(y[i], x[i]), 0, proton[i]};
}

void rotate_3d(MT64& rng, std::vector<Nucleon>& n) {
    const double al = 2 * M_PI * rng.real3();
    const double be = std::acos(1. - 2. * rng.real3());
    const double ga = 2 * M_PI * rng.real3();
    const double c1 = std::cos(al), s1 = std::sin(al), c2 = std::cos(be), s2 = std::sin(be);
    const double c3 = std::cos(ga), s3 = std::sin(ga);
    for (Nucleon& q : n) {
        const double x = q.x, y = q.y, z = q.z;
        q.x = c2 * x - c3 * s2 * y + s2 * s3 * z;
        q.y = c1 * s2 * x + (c1 * c2 * c3 - s1 * s3) * y + (-c3 * s1 - c1 * c2 * s3) * z;
        q.z = s1 * s2 * x + (c1 * s3 + c2 * c3 * s1) * y + (c1 * c3 - c2 * s1 * s3) * z;
    }
}

}  // namespace

// ---- deuteron --------------------------------------------------------------------------
namespace {

// Hulthen thickness integrand in xi = e^{-z a} (Glauber.cpp:874-903): the
// transverse distance s is in units of 1/a, the factor 2 folds z < 0
struct HulthenTA {
    double a, w, rho, sigma, s;
    double f(double xi) const {
        if (xi == 0.0) xi = 1e-10;
        if (xi == 1.0) xi = 1.0 - 1e-10;
        const double sa = s * a;
        const double z = -std::log(xi);
        const double r = std::sqrt(sa * sa + z * z);
        double v = 2.0 * a * rho * sigma;
        const double g = (1.0 / r) * (std::exp(-r) - std::exp(-(w / a) * r));
        v *= g * g;
        return v;
    }
};

// adaptive 7-point closed Newton-Cotes with interval halving
// (Glauber.cpp:905-1100): each level compares the halves against the
// parent estimate, tolerance divided by sqrt 2 per level, at most 10000
// integrand evaluations in all
double nc7(const HulthenTA& F, double tol, double lo, double dx, const double* fo, double pre, double area,
           int* count) {
    static const double wt[7] = {41.0 / 140.0, 54.0 / 35.0, 27.0 / 140.0, 68.0 / 35.0,
                                 27.0 / 140,   54.0 / 35.0, 41.0 / 140.0};
    dx /= 2.0;
    double fl[7], fr[7];
    fl[1] = F.f(lo + dx);
    fl[3] = F.f(lo + 3.0 * dx);
    fl[5] = F.f(lo + 5.0 * dx);
    fl[0] = fo[0];
    fl[2] = fo[1];
    fl[4] = fo[2];
    fl[6] = fo[3];
    *count += 3;
    double left = 0.0;
    for (int i = 0; i < 7; i++) left += wt[i] * fl[i];
    left *= dx;
    fr[1] = F.f(lo + 7.0 * dx);
    fr[3] = F.f(lo + 9.0 * dx);
    fr[5] = F.f(lo + 11.0 * dx);
    fr[0] = fo[3];
    fr[2] = fo[4];
    fr[4] = fo[5];
    fr[6] = fo[6];
    *count += 3;
    double right = 0.0;
    for (int i = 0; i < 7; i++) right += wt[i] * fr[i];
    right *= dx;
    double ans = left + right;
    area += -std::fabs(pre) + std::fabs(left) + std::fabs(right);
    if (std::fabs(ans - pre) > tol * std::fabs(area) && *count < 10000) {
        tol /= 1.414213562;
        left = nc7(F, tol, lo, dx, fl, left, area, count);
        right = nc7(F, tol, lo + dx * 6., dx, fr, right, area, count);
        ans = left + right;
    }
    return ans;
}

double nc7_integral(const HulthenTA& F) {
    const double lo = 0.0, dx = (1.0 - 0.0) / 6.0;
    double g[7];
    for (int i = 0; i < 7; i++) g[i] = F.f(lo + i * dx);
    int count = 7;
    return nc7(F, 1.0e-6, lo, dx, g, 0.0, 0.0, &count);
}

struct DeuteronTable {
    std::vector<double> vx, vy;
};

// 101 nodes on [0, 2 x 12 fm] (Glauber.cpp:400-430, initGlauber's 100
// intervals and SCutOff 12)
const DeuteronTable& deuteron_table(double a, double w, double R0, double sigma) {
    static std::mutex mu;
    static std::map<std::vector<double>, DeuteronTable> cache;
    std::lock_guard<std::mutex> lk(mu);
    const std::vector<double> key = {a, w, R0, sigma};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    // Glauber::CalcRho with AnumHulthen: int d^3r (e^{-ar} - e^{-wr})^2/r^2 =
    // 2 pi (a - w)^2 / (a w (a + w)); rho0 starts at R_WS (Glauber.cpp:246)
    const double f0 = 2.0 * (a - w) * (a - w) * M_PI / a / w / (a + w);
    const double rho_start = R0;
    const double f = (f0 * rho_start) / rho_start;
    HulthenTA F{a, w, 2.0 / f, sigma, 0.};
    DeuteronTable t;
    const int n = 100;
    const double up = 2.0 * 12.0, dx = (up - 0.0) / n;
    t.vx.resize(n + 1);
    t.vy.resize(n + 1);
    for (int i = 0; i <= n; i++) {
        t.vx[i] = dx * i;
        F.s = t.vx[i];
        t.vy[i] = nc7_integral(F);
    }
    return cache.emplace(key, std::move(t)).first->second;
}

}  // namespace

double deuteron_thickness(double s, double a, double w, double R0, double sigma) {
    const DeuteronTable& t = deuteron_table(a, w, R0, sigma);
    const int n = (int)t.vx.size() - 1;
    if (s > t.vx[n]) return 0.0;
    // cubic through four evenly spaced nodes, x between the second and third
    // (Glauber.cpp:300-358)
    int k = (int)(n * (s - t.vx[0]) / (t.vx[n] - t.vx[0])) - 1;
    if (k <= 0)
        k = 0;
    else if (k >= n - 3)
        k = n - 3;
    const double h = (t.vx[n] - t.vx[0]) / n;
    const double f0 = t.vy[k], f1 = t.vy[k + 1], f2 = t.vy[k + 2], f3 = t.vy[k + 3];
    const double c3 = (-f0 + 3.0 * f1 - 3.0 * f2 + f3) / (6.0 * h * h * h);
    const double c2 = (2.0 * f0 - 5.0 * f1 + 4.0 * f2 - f3) / (2.0 * h * h);
    const double c1 = (-11.0 * f0 + 18.0 * f1 - 9.0 * f2 + 2.0 * f3) / (6.0 * h);
    const double u = s - t.vx[k];
    double y = c3 * std::pow(u, 3.);
    y += c2 * std::pow(u, 2.);
    y += c1 * u;
    y += f0;
    return y < 0.0 ? 0.0 : y;
}

void sample_deuteron_separation(MT64& rng, double a, double w, double R0, double sigma, double* x, double* y) {
    // envelope A r 11 e^{-r^2/40}, scaled with sigma_NN against 42.13 mb;
    // r drawn from it by inverting its integral up to 15 fm
    const double A = 1.2 * sigma / 4.21325504715;
    const double area15 = A * 220. * (1. - std::exp(-0.025 * 15. * 15.));
    double phi, r, u;
    do {
        phi = 2. * M_PI * rng.real1();
        r = 6.32456 * std::sqrt(-std::log((-0.00454545 * (-220. * A + area15 * rng.real1())) / A));
        u = rng.real1();
    } while (u > r * deuteron_thickness(r, a, w, R0, sigma) / (A * r * 11. * std::exp(-r * r / 40.)));
    *x = r * std::cos(phi);
    *y = r * std::sin(phi);
}

int write_npart_dndy(const char* path, int Npart, double dNdy, double Tpp, double b, double dEdy,
                     unsigned long long seed, double n3, double e3, double n6, double e6, double g, double muZero,
                     double c, double fQs, double avgQs, std::string& err) {
    std::ofstream out(path, std::ios::out);
    if (!out) {
        err = std::string("cannot write ") + path;
        return IPG_ERR_IO;
    }
    const double last =
        g * g /
        (4. * M_PI * 4. * M_PI /
         (9. * std::log(std::pow(std::pow(muZero / 0.2, 2. / c) + std::pow(fQs * avgQs / 0.2, 2. / c), c))));
    out << Npart << " " << dNdy << " " << Tpp << " " << b << " " << dEdy << " " << seed << " N/A N/A N/A " << n3
        << " " << e3 << " " << n6 << " " << e6 << " " << last << std::endl;
    return IPG_OK;
}

// the first draw of Init::init (Init.cpp:2320-2347)
double sample_b(MT64& rng, const ipg_params& p) {
    const double xb = rng.real1();
    return p.samplebFromLinearDistribution == 1 ? std::sqrt((p.bmax * p.bmax - p.bmin * p.bmin) * xb + p.bmin * p.bmin)
                                                : (p.bmax - p.bmin) * xb + p.bmin;
}

int sample_event(MT64& rng, const ipg_params& p, const NucleusData& proj, const NucleusData& targ,
                 std::vector<Nucleon>& A, std::vector<Nucleon>& B, EventInfo& info, std::string& err) {
    if (p.nucleonPositionsFromFile != 0) {
        err = "nucleonPositionsFromFile != 0 is not supported";
        return IPG_ERR_ARG;
    }
    if (p.useConstituentQuarkProton > 0 || p.smearQs == 1 || p.averageOverThisManyNuclei != 1) {
        err = "constituent quarks, smearQs and nucleus averaging are not supported";
        return IPG_ERR_ARG;
    }
    if (proj.A == 3 || targ.A == 3) {
        err = "the He3 sampler is not supported";
        return IPG_ERR_ARG;
    }
    const double b = sample_b(rng, p);
    A.clear();
    B.clear();
    const double sigma_fm2 = 0.1 * p.SigmaNN;
    // the deuteron (Init.cpp:90-113, 129-152): p-n separation, nucleons at
    // +-1/2 of it; the second nucleon keeps proton = 1 in the projectile and
    // gets proton = 0 in the target, as the reference writes them
    auto deuteron = [&](const NucleusData& d, std::vector<Nucleon>& out, int second_proton) {
        double x, y;
        sample_deuteron_separation(rng, d.a, d.w, d.R, sigma_fm2, &x, &y);
        out.push_back(Nucleon{x / 2., y / 2., 0., 0., 0, 1});
        out.push_back(Nucleon{-(x / 2.), -(y / 2.), -0., 0., 0, second_proton});
    };
    if (proj.A == 1)
        A.push_back(Nucleon{0., 0., 0., 0., 0, 1});
    else if (proj.A == 2)
        deuteron(proj, A, 1);
    else
        generate_nucleus(rng, proj, A);
    if (targ.A == 1)
        B.push_back(Nucleon{0., 0., 0., 0., 0, 1});
    else if (targ.A == 2)
        deuteron(targ, B, 0);
    else
        generate_nucleus(rng, targ, B);
    rotate_3d(rng, A);
    rotate_3d(rng, B);

    double phiRP = 0.;
    if (p.rotateReactionPlane) phiRP = 2 * M_PI * rng.real2();
    for (Nucleon& q : A) {
        q.x -= b / 2. * std::cos(phiRP);
        q.y -= b / 2. * std::sin(phiRP);
    }
    for (Nucleon& q : B) {
        q.x += b / 2. * std::cos(phiRP);
        q.y += b / 2. * std::sin(phiRP);
    }
    const double xi = p.protonAnisotropy;
    if (xi != 0.) {
        for (Nucleon& q : A) q.phi = 2 * M_PI * rng.real2();
        for (Nucleon& q : B) q.phi = 2 * M_PI * rng.real2();
    } else {
        for (Nucleon& q : A) q.phi = 0.;
        for (Nucleon& q : B) q.phi = 0.;
    }
    // binary collisions (Init.cpp:1117-1216)
    const double d2 = p.SigmaNN / (M_PI * 10.);
    int Ncoll = 0;
    for (size_t i = 0; i < A.size(); i++) {
        for (size_t j = 0; j < B.size(); j++) {
            const double dx = B[j].x - A[i].x, dy = B[j].y - A[i].y;
            const double dij = dx * dx + dy * dy;
            bool hit;
            if (p.gaussianWounding == 0) {
                hit = dij < d2;
            } else {
                const double G = 0.92;
                const double pr = G * std::exp(-G * dij / d2);
                hit = rng.real1() < pr;
            }
            if (hit) {
                Ncoll++;
                B[j].collided = 1;
                A[i].collided = 1;
            }
        }
    }
    if (A.size() == 1 && B.size() == 1) {
        A[0].collided = 1;
        B[0].collided = 1;
    }
    int Npart = 0;
    for (const Nucleon& q : A) Npart += q.collided == 1;
    for (const Nucleon& q : B) Npart += q.collided == 1;
    info.b = b;
    info.phiRP = phiRP;
    info.Ncoll = Ncoll;
    info.Npart = Npart;
    return IPG_OK;
}

// ---- input file -----------------------------------------------------------------
void default_params(ipg_params* p) {
    std::memset(p, 0, sizeof(*p));
    // the reference's shipped `input` file
    p->size = 64;
    p->Nc = 3;
    p->Ny = 50;
    p->L = 8;
    p->g = 1.;
    p->m = 0.4;
    p->UVdamp = 0.;
    p->BG = 4.;
    p->QsmuRatio = 0.8;
    p->rmax = 1e7;
    p->LOutput = 30.;
    p->sizeOutput = 512;
    p->etaSizeOutput = 1;
    p->detaOutput = 0.;
    p->roots = 7000;
    p->Rapidity = 0.;
    p->xFromThisFactorTimesQs = 1;
    p->SigmaNN = 73.;
    p->maxtime = 0.1;
    p->useFluctuatingx = 1;
    p->useNucleus = 1;
    p->gaussianWounding = 1;
    p->rotateReactionPlane = 1;
    p->samplebFromLinearDistribution = 1;
    p->averageOverThisManyNuclei = 1;
    p->useTimeForSeed = 1;
    p->computeGluonMultiplicity = 1;
    p->R_WS = 6.6;
    p->a_WS = 0.52;
    p->beta2 = 0.28;
    p->beta4 = 0.08;
    p->d_min = 0.9;
    p->bmin = 0;
    p->bmax = 5;
    p->smearingWidth = 0.6;
    p->Jacobianm = 0.35;
    p->g2mu = 0.1;
    p->muZero = 0.3;
    p->c = 0.2;
    p->runWithThisFactorTimesQs = 0.5;
    p->seed = 13;
    std::strcpy(p->Projectile, "p");
    std::strcpy(p->Target, "p");
    std::strcpy(p->NucleusQsTableFileName, "qs2Adj_vs_Tp_vs_Y_200.in");
    derive(p);
}

void derive(ipg_params* p) {
    const double a = p->L / static_cast<double>(p->size);
    const int steps = static_cast<int>(10 * p->maxtime / a) + 1;
    p->dtau = p->maxtime / (steps * a);
    p->itmax = static_cast<int>(p->maxtime / (a * p->dtau) + 0.1);
}

int read_input(const char* path, ipg_params* p, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = std::string("input file ") + path + " is absent";
        return IPG_ERR_IO;
    }
    // pairwise scan up to EndOfFile; the first occurrence of a key wins
    std::map<std::string, std::string> kv;
    std::string k, v;
    while (in >> k) {
        if (k == "EndOfFile") break;
        if (!(in >> v)) break;
        if (!kv.count(k)) kv[k] = v;
    }
    auto has = [&](const char* key) { return kv.count(key) > 0; };
    auto D = [&](const char* key, double& dst) {
        if (has(key)) dst = std::atof(kv[key].c_str());
    };
    auto I = [&](const char* key, int& dst) {
        if (has(key)) dst = static_cast<int>(std::atof(kv[key].c_str()));
    };
    auto Sx = [&](const char* key, char* dst, size_t n) {
        if (has(key)) {
            std::strncpy(dst, kv[key].c_str(), n - 1);
            dst[n - 1] = 0;
        }
    };
    Sx("NucleusQsTableFileName", p->NucleusQsTableFileName, sizeof(p->NucleusQsTableFileName));
    I("nucleonPositionsFromFile", p->nucleonPositionsFromFile);
    Sx("Target", p->Target, sizeof(p->Target));
    Sx("Projectile", p->Projectile, sizeof(p->Projectile));
    I("runningCoupling", p->runningCoupling);
    D("L", p->L);
    D("BG", p->BG);
    I("size", p->size);
    I("useFluctuatingx", p->useFluctuatingx);
    I("Nc", p->Nc);
    I("inverseQsForMaxTime", p->inverseQsForMaxTime);
    if (has("seed")) p->seed = (unsigned long long)(std::atof(kv["seed"].c_str()) + 0.5);
    I("useSeedList", p->useSeedList);
    I("Ny", p->Ny);
    D("roots", p->roots);
    D("muZero", p->muZero);
    D("c", p->c);
    D("runWithThisFactorTimesQs", p->runWithThisFactorTimesQs);
    D("g", p->g);
    D("m", p->m);
    D("Jacobianm", p->Jacobianm);
    D("SigmaNN", p->SigmaNN);
    D("rmax", p->rmax);
    D("UVdamp", p->UVdamp);
    I("setWSDeformParams", p->setWSDeformParams);
    if (p->setWSDeformParams) {
        D("R_WS", p->R_WS);
        D("a_WS", p->a_WS);
        D("beta2", p->beta2);
        D("beta3", p->beta3);
        D("beta4", p->beta4);
        D("gamma", p->gamma);
        double f = p->force_dmin_flag;
        D("force_dmin_flag", f);
        p->force_dmin_flag = f != 0.;
        D("d_min", p->d_min);
        D("dR_np", p->dR_np);
        D("da_np", p->da_np);
    }
    D("bmin", p->bmin);
    D("bmax", p->bmax);
    I("rotateReactionPlane", p->rotateReactionPlane);
    I("computeGluonMultiplicity", p->computeGluonMultiplicity);
    D("QsmuRatio", p->QsmuRatio);
    {
        double f = p->usePseudoRapidity;
        D("usePseudoRapidity", f);
        p->usePseudoRapidity = (int)f;
    }
    D("Rapidity", p->Rapidity);
    I("useNucleus", p->useNucleus);
    D("g2mu", p->g2mu);
    D("maxtime", p->maxtime);
    D("xFromThisFactorTimesQs", p->xFromThisFactorTimesQs);
    I("samplebFromLinearDistribution", p->samplebFromLinearDistribution);
    I("writeOutputs", p->writeOutputs);
    I("writeInitialWilsonLines", p->writeInitialWilsonLines);
    I("readInitialWilsonLines", p->readInitialWilsonLines);
    I("averageOverThisManyNuclei", p->averageOverThisManyNuclei);
    I("useTimeForSeed", p->useTimeForSeed);
    I("useFixedNpart", p->useFixedNpart);
    I("smearQs", p->smearQs);
    D("smearingWidth", p->smearingWidth);
    I("gaussianWounding", p->gaussianWounding);
    D("protonAnisotropy", p->protonAnisotropy);
    {
        double f = p->useConstituentQuarkProton;
        D("useConstituentQuarkProton", f);
        p->useConstituentQuarkProton = (int)f;
    }
    I("useSmoothNucleus", p->useSmoothNucleus);
    {
        // Setup::IFind truncates, and the reference stores it as an int
        int q = (int)p->minimumQs2ST;
        I("minimumQs2ST", q);
        p->minimumQs2ST = q;
    }
    I("SubNucleonParamType", p->SubNucleonParamType);
    I("rho_rng", p->rho_rng);
    I("newton_polish", p->newton_polish);
    I("sizeOutput", p->sizeOutput);
    I("etaSizeOutput", p->etaSizeOutput);
    D("LOutput", p->LOutput);
    D("detaOutput", p->detaOutput);
    derive(p);
    return IPG_OK;
}


// ---- Wilson-line files (Init.cpp:1857-1980 writer, 1988-2301 reader) ----------
//
// Text (format 1): one line per site in (x, y) order, "ix iy" and the 18
// numbers of the matrix COLUMN by column (Matrix::MatrixToString,
// Matrix.cpp:606-627), printed as an ostream with precision 15 prints them
// (%.15g); an empty line after each x row.  The text reader takes the 18
// numbers ROW by row, so a text file written and read back gives V^T, in the
// reference as here.
// Binary (format 2): int N, int Nc, double L, double a, double rapidity,
// then for x-major file position (p, q) = (slow, fast) the site with
// i = q, j = p, entries row by row as (re, im) doubles.

int write_wilson_lines(const char* path, int N, double L, double rapidity, const double* V, int format,
                       std::string& err) {
    if (format != 1 && format != 2) {
        err = "unknown Wilson-line format " + std::to_string(format) + " (1 = text, 2 = binary)";
        return IPG_ERR_ARG;
    }
    FILE* f = std::fopen(path, format == 1 ? "w" : "wb");
    if (!f) {
        err = std::string("cannot open ") + path + " for writing";
        return IPG_ERR_IO;
    }
    bool ok = true;
    if (format == 1) {
        for (int i = 0; i < N && ok; i++) {
            for (int j = 0; j < N && ok; j++) {
                const double* m = V + 18 * ((size_t)i * N + j);
                ok = std::fprintf(f, "%d %d", i, j) > 0;
                for (int c = 0; c < 3; c++)
                    for (int r = 0; r < 3; r++)
                        ok = ok && std::fprintf(f, " %.15g %.15g", m[2 * (3 * r + c)], m[2 * (3 * r + c) + 1]) > 0;
                ok = ok && std::fputc('\n', f) != EOF;
            }
            ok = ok && std::fputc('\n', f) != EOF;
        }
    } else {
        const int nc = 3;
        const double a = L / N;
        ok = std::fwrite(&N, sizeof(int), 1, f) == 1 && std::fwrite(&nc, sizeof(int), 1, f) == 1 &&
             std::fwrite(&L, sizeof(double), 1, f) == 1 && std::fwrite(&a, sizeof(double), 1, f) == 1 &&
             std::fwrite(&rapidity, sizeof(double), 1, f) == 1;
        std::vector<double> row(18 * (size_t)N);
        for (int p = 0; p < N && ok; p++) {
            for (int q = 0; q < N; q++) std::memcpy(&row[18 * (size_t)q], V + 18 * ((size_t)q * N + p), 18 * sizeof(double));
            ok = std::fwrite(row.data(), sizeof(double), row.size(), f) == row.size();
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        err = std::string("writing ") + path + " failed";
        return IPG_ERR_IO;
    }
    return IPG_OK;
}

int read_wilson_lines(const char* path, int N, double L, double b, int nucleus, int format, double* V,
                      std::string& err) {
    if (format != 1 && format != 2) {
        err = "unknown Wilson-line format " + std::to_string(format) + " (1 = text, 2 = binary)";
        return IPG_ERR_ARG;
    }
    if (nucleus != 0 && nucleus != 1) {
        err = "nucleus must be 0 (A) or 1 (B)";
        return IPG_ERR_ARG;
    }
    const double shift = nucleus == 0 ? -b / 2. : b / 2.;
    const size_t S = (size_t)N * N;
    if (format == 1) {
        std::ifstream in(path);
        if (!in) {
            err = std::string("Wilson-line file ") + path + " not found";
            return IPG_ERR_IO;
        }
        double re[9] = {0}, im[9] = {0}, skip = 0.;
        const double a = L / static_cast<double>(N);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                in >> skip >> skip;
                for (int k = 0; k < 9; k++) in >> re[k] >> im[k];
                // the row a line lands on: truncation toward zero, as the
                // reference converts (a i -+ b/2) / a
                const double x = a * i + shift;
                const int ix = static_cast<int>(x / a);
                if (ix < 0 || ix >= N) continue;
                double* m = V + 18 * ((size_t)ix * N + j);
                for (int k = 0; k < 9; k++) {
                    m[2 * k] = re[k];
                    m[2 * k + 1] = im[k];
                }
            }
        return IPG_OK;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in.good()) {
        err = std::string("Wilson-line file ") + path + " does not exist";
        return IPG_ERR_IO;
    }
    int n = 0, nc = 0;
    double Lf = 0., af = 0., rap = 0.;
    in.read(reinterpret_cast<char*>(&n), sizeof(int));
    in.read(reinterpret_cast<char*>(&nc), sizeof(int));
    in.read(reinterpret_cast<char*>(&Lf), sizeof(double));
    in.read(reinterpret_cast<char*>(&af), sizeof(double));
    in.read(reinterpret_cast<char*>(&rap), sizeof(double));
    if (!in) {
        err = std::string(path) + ": truncated header";
        return IPG_ERR_IO;
    }
    if (n != N) {
        err = std::string(path) + ": wrong lattice size, data is " + std::to_string(n) + " but the run has " +
              std::to_string(N);
        return IPG_ERR_ARG;
    }
    if (std::abs(Lf - L) > 1e-5) {
        err = std::string(path) + ": grid length " + std::to_string(Lf) + " but the run has " + std::to_string(L);
        return IPG_ERR_ARG;
    }
    if (nc != 3) {
        err = std::string(path) + ": Nc = " + std::to_string(nc) + ", only 3 is supported";
        return IPG_ERR_ARG;
    }
    // values in file order; the spacing is the file's L over N
    const double a = Lf / static_cast<double>(N);
    double v[2];
    for (long long t = 0; in.read(reinterpret_cast<char*>(v), sizeof(v)); t++) {
        const long long pos = t / 9;
        const int e = (int)(t - pos * 9);
        const long long slow = pos / N, fast = pos - slow * N;
        const int ix = (int)std::round((a * (double)fast + shift) / a);
        const long long site = (long long)N * ix + slow;
        if (site < 0 || site >= (long long)S) continue;
        V[18 * (size_t)site + 2 * e] = v[0];
        V[18 * (size_t)site + 2 * e + 1] = v[1];
    }
    return IPG_OK;
}


// ---- eccentricities<id>.dat (Evolution.cpp:1280-1810) --------------------------

void ecc_extents(const double* eps, const double* utau, int N, double L, double avx, double avy, double cutoff,
                 double* maxX, double* maxY) {
    const double a = L / N;
    const int xshift = static_cast<int>(std::floor(avx / a + 0.00000000001));
    const int yshift = static_cast<int>(std::floor(avy / a + 0.00000000001));
    double mx = 0., my = 0.;
    for (int ix = 2; ix < N - 2; ix++) {
        const double x = -L / 2. + a * ix - avx;
        for (int iy = 2; iy < N - 2; iy++) {
            const double y = -L / 2. + a * iy - avy;
            const size_t pos = (size_t)ix * N + iy;
            const double w = eps[pos] < cutoff ? 0. : eps[pos] * utau[pos];
            if (w > cutoff && iy == N / 2 + yshift) mx = x;
            if (w > cutoff && ix == N / 2 + xshift) my = y;
        }
    }
    *maxX = mx;
    *maxY = my;
}

int append_eccentricity_line(const char* path, double tau, const double* ecc, const double* psi, double cutoff,
                             double rms_r, double maxX, double maxY, double b, double Tpp, double area, double Rbar,
                             double avg_eps, double avg_qs, std::string& err) {
    FILE* f = std::fopen(path, "a");
    if (!f) {
        err = std::string("cannot open ") + path + " for appending";
        return IPG_ERR_IO;
    }
    // an ostream's default format: %g, six significant digits
    bool ok = std::fprintf(f, "%g", tau) > 0;
    for (int n = 0; n < 6; n++) ok = ok && std::fprintf(f, " %g %g", ecc[n], psi[n]) > 0;
    const double rest[10] = {cutoff, rms_r, maxX, maxY, b, Tpp, area, Rbar, avg_eps, avg_qs};
    for (double v : rest) ok = ok && std::fprintf(f, " %g", v) > 0;
    ok = ok && std::fputc('\n', f) != EOF;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        err = std::string("writing ") + path + " failed";
        return IPG_ERR_IO;
    }
    return IPG_OK;
}


// ---- epsilon-u-Hydro output (MyEigen.cpp:326-800) --------------------------------
//
// Each hydro cell (x, y) inside the lattice takes the four lattice sites
// around it, xpos = floor((x + L/2)/a + 1e-10) and the next one (clamped at
// N-1), and interpolates each quantity q as
//   x1 = (1 - fx) q(xpos, ypos) + fx q(xpos+1, ypos)    (0 unless both sites
//   x2 = (1 - fx) q(xpos, ypos+1) + fx q(xpos+1, ypos+1)  are in (0, N^2))
//   q  = (1 - fy) x1 + fy x2,   fx = (x - x_low) / ha,  fy = (y - y_low) / ha
// with ha the hydro spacing (the reference divides by ha, not by a), |eps|
// for the energy density; u^tau is then rebuilt from u^x, u^y, u^eta.  The
// columns: eta x y eps[GeV/fm^3] u^tau u^x u^y u^eta pi^{tt tx ty teta xx xy
// xeta yy yeta etaeta}, a vacuum line "eta x y 0 1 0 ..." where eps <= 1e-10
// or outside the lattice, an empty line after each eta slice.

int write_hydro(const char* path, const double* land, int N, double L, int it, double dtau, int hx, int heta,
                double hL, double deta, double* Etot, std::string& err) {
    FILE* f = std::fopen(path, "w");
    if (!f) {
        err = std::string("cannot open ") + path + " for writing";
        return IPG_ERR_IO;
    }
    const size_t S = (size_t)N * N;
    const double a = L / N;
    const int hy = hx;
    const double tau0 = it * dtau * a;
    const double ha = hL / static_cast<double>(hx);
    // land planes in the order of the columns after u^eta
    const int pic[10] = {5, 9, 10, 11, 6, 12, 13, 7, 14, 8};
    bool ok = std::fprintf(f, "# dummy %d etamax= %d xmax= %d ymax= %d deta= %g dx= %g dy= %g tau= %g\n", 1, heta, hx,
                           hy, deta, ha, ha, tau0) > 0;
    double etot = 0.;
    for (int ieta = 0; ieta < heta && ok; ieta++) {
        const double eta = -(heta - 1) / 2. * deta + deta * ieta;
        for (int ix = 0; ix < hx && ok; ix++) {
            for (int iy = 0; iy < hy && ok; iy++) {
                const double x = -hL / 2. + ha * ix;
                const double y = -hL / 2. + ha * iy;
                double v[15];
                bool inside = std::abs(x) < L / 2. && std::abs(y) < L / 2.;
                if (inside) {
                    const int xpos = static_cast<int>(std::floor((x + L / 2.) / a + 0.0000000001));
                    const int ypos = static_cast<int>(std::floor((y + L / 2.) / a + 0.0000000001));
                    const int xup = xpos < N - 1 ? xpos + 1 : xpos;
                    const int yup = ypos < N - 1 ? ypos + 1 : ypos;
                    const double xlow = -L / 2. + a * xpos, ylow = -L / 2. + a * ypos;
                    const double fracx = (x - xlow) / ha;
                    const double fracy = (y - ylow) / ha;
                    const long long p1 = (long long)xpos * N + ypos, p2 = (long long)xup * N + ypos;
                    const long long p3 = (long long)xpos * N + yup, p4 = (long long)xup * N + yup;
                    const long long SS = (long long)S;
                    const bool lo = p1 > 0 && p1 < SS && p2 > 0 && p2 < SS;
                    const bool hi = p3 > 0 && p3 < SS && p4 > 0 && p4 < SS;
                    for (int q = 0; q < 15; q++) {
                        const double* F = land + (size_t)q * S;
                        double f1 = 0., f2 = 0., f3 = 0., f4 = 0.;
                        if (lo) f1 = F[p1], f2 = F[p2];
                        if (hi) f3 = F[p3], f4 = F[p4];
                        if (q == 0) f1 = std::abs(f1), f2 = std::abs(f2), f3 = std::abs(f3), f4 = std::abs(f4);
                        const double x1 = lo ? (1 - fracx) * f1 + fracx * f2 : 0.;
                        const double x2 = hi ? (1 - fracx) * f3 + fracx * f4 : 0.;
                        v[q] = (1. - fracy) * x1 + fracy * x2;
                    }
                    v[1] = std::sqrt(1. + v[2] * v[2] + v[3] * v[3] + tau0 * tau0 * v[4] * v[4]);
                    const double e = std::abs(HBARC * v[0] * 1.);
                    etot += e * ha * ha * it * dtau * a;
                    inside = e > 0.0000000001;
                    if (inside) {
                        ok = std::fprintf(f, "%g %g %g %g %g %g %g %g", eta, x, y, e, v[1], v[2], v[3], v[4]) > 0;
                        for (int k = 0; k < 10; k++) ok = ok && std::fprintf(f, " %g", v[pic[k]] * 1.) > 0;
                        ok = ok && std::fputc('\n', f) != EOF;
                    }
                }
                if (!inside) {
                    ok = std::fprintf(f, "%g %g %g 0 1 0 0 0 0 0 0 0 0 0 0 0 0 0\n", eta, x, y) > 0;
                }
            }
        }
        ok = ok && std::fputc('\n', f) != EOF;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        err = std::string("writing ") + path + " failed";
        return IPG_ERR_IO;
    }
    if (Etot) *Etot = etot;
    return IPG_OK;
}


// ---- Tmunu-t output (MyEigen.cpp:801-1046) -----------------------------------------
//
// The same interpolation as the hydro file (without |.|) for the ten
// components, y rows outer and x inner, an empty line after each row.
// Columns: ix iy, then in GeV/fm^3 T^tt, T^xx, T^yy, tau^2 T^etaeta, -T^tx,
// -T^ty, -tau T^teta, -T^xy, -tau T^yeta, -tau T^xeta; a floor line
// "ix iy 1e-16 5e-17 5e-17 0 ..." where T^tt hbarc <= 1e-16 or outside the
// lattice (small_eps, Phys_consts.h:5).

int write_tmunu(const char* path, const double* T, int N, double L, int it, double dtau, int hx, int heta, double hL,
                double deta, std::string& err) {
    FILE* f = std::fopen(path, "w");
    if (!f) {
        err = std::string("cannot open ") + path + " for writing";
        return IPG_ERR_IO;
    }
    const double small_eps = 1e-16;
    const long long S = (long long)N * N;
    const double a = L / N;
    const int hy = hx;
    const double tau0 = it * dtau * a;
    const double ha = hL / static_cast<double>(hx);
    bool ok = std::fprintf(f, "# dummy %d etamax= %d xmax= %d ymax= %d deta= %g dx= %g dy= %g\n", 1, heta, hx, hy,
                           deta, ha, ha) > 0;
    for (int iy = 0; iy < hy && ok; iy++) {
        for (int ix = 0; ix < hx && ok; ix++) {
            const double x = -hL / 2. + ha * ix;
            const double y = -hL / 2. + ha * iy;
            bool written = false;
            if (std::abs(x) < L / 2. && std::abs(y) < L / 2.) {
                const int xpos = static_cast<int>(std::floor((x + L / 2.) / a + 0.0000000001));
                const int ypos = static_cast<int>(std::floor((y + L / 2.) / a + 0.0000000001));
                const int xup = xpos < N - 1 ? xpos + 1 : xpos;
                const int yup = ypos < N - 1 ? ypos + 1 : ypos;
                const double xlow = -L / 2. + a * xpos, ylow = -L / 2. + a * ypos;
                const long long p1 = (long long)xpos * N + ypos, p2 = (long long)xup * N + ypos;
                const long long p3 = (long long)xpos * N + yup, p4 = (long long)xup * N + yup;
                const double fracx = (x - xlow) / ha;
                const double fracy = (y - ylow) / ha;
                const bool lo = p1 > 0 && p1 < S && p2 > 0 && p2 < S;
                const bool hi = p3 > 0 && p3 < S && p4 > 0 && p4 < S;
                double r[10];
                for (int q = 0; q < 10; q++) {
                    const double* F = T + (size_t)q * S;
                    const double x1 = lo ? (1 - fracx) * F[p1] + fracx * F[p2] : 0.;
                    const double x2 = hi ? (1 - fracx) * F[p3] + fracx * F[p4] : 0.;
                    r[q] = (1. - fracy) * x1 + fracy * x2;
                }
                // T's order: tt 0, xx 1, yy 2, ee 3, tx 4, ty 5, te 6, xy 7, xe 8, ye 9
                if (r[0] * 1. * HBARC > small_eps) {
                    ok = std::fprintf(f, "%d %d %g %g %g %g %g %g %g %g %g %g\n", ix, iy, r[0] * 1. * HBARC,
                                      r[1] * 1. * HBARC, r[2] * 1. * HBARC, tau0 * tau0 * r[3] * 1. * HBARC,
                                      -r[4] * 1. * HBARC, -r[5] * 1. * HBARC, -tau0 * r[6] * 1. * HBARC,
                                      -r[7] * 1. * HBARC, -tau0 * r[9] * 1. * HBARC, -tau0 * r[8] * 1. * HBARC) > 0;
                    written = true;
                }
            }
            if (!written && ok)
                ok = std::fprintf(f, "%d %d %g %g %g 0 0 0 0 0 0 0\n", ix, iy, small_eps, small_eps / 2.,
                                  small_eps / 2.) > 0;
        }
        ok = ok && std::fputc('\n', f) != EOF;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        err = std::string("writing ") + path + " failed";
        return IPG_ERR_IO;
    }
    return IPG_OK;
}


// ---- Jazma-Hydro output (MyEigen.cpp:1047-1293) ---------------------------------
//
// Two passes over the hydro grid with the hydro file's interpolation of
// |g2mu2A| and |g2mu2B|: the sum J of g2mu2A g2mu2B ha^2 tau over all eta
// slices, then one line per cell, "eta x y (g2mu2A g2mu2B / J) Etot 1 0 ...",
// the vacuum line outside the lattice, no empty lines.

namespace {
struct HydroCell {
    bool inside = false, lo = false, hi = false;
    double fx = 0., fy = 0.;
    long long p1 = 0, p2 = 0, p3 = 0, p4 = 0;
};
HydroCell hydro_cell(double x, double y, int N, double L, double ha) {
    HydroCell c;
    c.inside = std::abs(x) < L / 2. && std::abs(y) < L / 2.;
    if (!c.inside) return c;
    const double a = L / N;
    const long long S = (long long)N * N;
    const int xpos = static_cast<int>(std::floor((x + L / 2.) / a + 0.0000000001));
    const int ypos = static_cast<int>(std::floor((y + L / 2.) / a + 0.0000000001));
    const int xup = xpos < N - 1 ? xpos + 1 : xpos;
    const int yup = ypos < N - 1 ? ypos + 1 : ypos;
    c.fx = (x - (-L / 2. + a * xpos)) / ha;
    c.fy = (y - (-L / 2. + a * ypos)) / ha;
    c.p1 = (long long)xpos * N + ypos;
    c.p2 = (long long)xup * N + ypos;
    c.p3 = (long long)xpos * N + yup;
    c.p4 = (long long)xup * N + yup;
    c.lo = c.p1 > 0 && c.p1 < S && c.p2 > 0 && c.p2 < S;
    c.hi = c.p3 > 0 && c.p3 < S && c.p4 > 0 && c.p4 < S;
    return c;
}
double interp_abs(const HydroCell& c, const double* F) {
    const double x1 = c.lo ? (1 - c.fx) * std::abs(F[c.p1]) + c.fx * std::abs(F[c.p2]) : 0.;
    const double x2 = c.hi ? (1 - c.fx) * std::abs(F[c.p3]) + c.fx * std::abs(F[c.p4]) : 0.;
    return (1. - c.fy) * x1 + c.fy * x2;
}
}  // namespace

int write_jazma(const char* path, const double* g2A, const double* g2B, int N, double L, int it, double dtau, int hx,
                int heta, double hL, double deta, double Etot, std::string& err) {
    const double a = L / N;
    const int hy = hx;
    const double ha = hL / static_cast<double>(hx);
    double J = 0.;
    for (int ieta = 0; ieta < heta; ieta++)
        for (int ix = 0; ix < hx; ix++)
            for (int iy = 0; iy < hy; iy++) {
                const HydroCell c = hydro_cell(-hL / 2. + ha * ix, -hL / 2. + ha * iy, N, L, ha);
                if (c.inside) J += interp_abs(c, g2A) * interp_abs(c, g2B) * ha * ha * it * dtau * a;
            }
    FILE* f = std::fopen(path, "w");
    if (!f) {
        err = std::string("cannot open ") + path + " for writing";
        return IPG_ERR_IO;
    }
    bool ok = std::fprintf(f, "# dummy %d etamax= %d xmax= %d ymax= %d deta= %g dx= %g dy= %g\n", 1, heta, hx, hy,
                           deta, ha, ha) > 0;
    for (int ieta = 0; ieta < heta && ok; ieta++) {
        const double eta = -(heta - 1) / 2. * deta + deta * ieta;
        for (int ix = 0; ix < hx && ok; ix++)
            for (int iy = 0; iy < hy && ok; iy++) {
                const double x = -hL / 2. + ha * ix, y = -hL / 2. + ha * iy;
                const HydroCell c = hydro_cell(x, y, N, L, ha);
                const double e = c.inside ? interp_abs(c, g2A) * interp_abs(c, g2B) / J * Etot : 0.;
                ok = std::fprintf(f, "%g %g %g %g 1 0 0 0 0 0 0 0 0 0 0 0 0 0\n", eta, x, y, e) > 0;
            }
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        err = std::string("writing ") + path + " failed";
        return IPG_ERR_IO;
    }
    return IPG_OK;
}

}  // namespace ipg
