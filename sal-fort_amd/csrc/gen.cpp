// Seeded synthetic inputs: grids of unit vectors with cell areas, and sea-surface-height fields.
//
// The reference reads its grid from CSV files that are not distributed (code.f90:36-56); these
// generators stand in for them.  Random numbers come from a stateless counter generator
// (SplitMix64 finaliser applied to (seed, stream, counter)), so every value depends only on its
// index and the threaded loops are deterministic for any thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"

namespace sal {
namespace {

inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t ctr) { return mix64(mix64(mix64(seed) ^ stream) ^ ctr); }
inline double uni(uint64_t seed, uint64_t stream, uint64_t ctr) {  // [0, 1)
    return (double)(rnd(seed, stream, ctr) >> 11) * 0x1.0p-53;
}
inline double gauss(uint64_t seed, uint64_t stream, uint64_t ctr) {  // Box-Muller on two counters
    double u1 = 1.0 - uni(seed, stream, 2 * ctr), u2 = uni(seed, stream, 2 * ctr + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

template <class F>
void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&f, a, b] {
            for (int64_t i = a; i < b; ++i) f(i);
        });
    }
    for (auto& t : th) t.join();
}

// fully normalised real spherical harmonics: y[l*l + l + m], m in [-l, l], orthonormal on the sphere
void real_sh(int L, double x, double y, double z, double* Y) {
    double st = std::sqrt(std::max(0.0, 1.0 - z * z));
    double cp = 1.0, sp = 0.0;
    double rxy = std::sqrt(x * x + y * y);
    if (rxy > 0) { cp = x / rxy; sp = y / rxy; }
    double pmm = std::sqrt(1.0 / (4.0 * M_PI));
    double cm = 1.0, sm = 0.0;  // cos(m phi), sin(m phi)
    for (int m = 0; m <= L; ++m) {
        if (m > 0) {
            pmm *= std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * st;
            double c2 = cm * cp - sm * sp, s2 = sm * cp + cm * sp;
            cm = c2;
            sm = s2;
        }
        double p2 = 0.0, p1 = pmm;  // P(l-2), P(l-1)
        for (int l = m; l <= L; ++l) {
            double p;
            if (l == m) p = pmm;
            else if (l == m + 1) p = std::sqrt(2.0 * m + 3.0) * z * pmm;
            else {
                double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                p = a * (z * p1 - b * p2);
            }
            if (l > m) { p2 = p1; p1 = p; }
            if (m == 0) Y[l * l + l] = p;
            else {
                Y[l * l + l + m] = M_SQRT2 * p * cm;
                Y[l * l + l - m] = M_SQRT2 * p * sm;
            }
        }
    }
}

}  // namespace
}  // namespace sal

using namespace sal;

extern "C" {

int sal_gen_fibonacci(int64_t n, double* x, double* y, double* z) {
    return guard([&] {
        SAL_REQUIRE(n >= 1 && x && y && z, SAL_ERR_ARG, "bad arguments");
        const double ga = M_PI * (3.0 - std::sqrt(5.0));
        parallel_for(n, [&](int64_t i) {
            double zz = 1.0 - (2.0 * i + 1.0) / (double)n;
            double r = std::sqrt(1.0 - zz * zz), ph = (double)i * ga;
            x[i] = r * std::cos(ph);
            y[i] = r * std::sin(ph);
            z[i] = zz;
        });
    });
}

int sal_gen_mpas(int64_t n, uint64_t seed, double total_area, double* x, double* y, double* z, double* area) {
    return guard([&] {
        SAL_REQUIRE(n >= 1 && x && y && z && area, SAL_ERR_ARG, "bad arguments");
        const double lat0 = 40.0 * M_PI / 180.0, lon0 = -60.0 * M_PI / 180.0, sig = 15.0 * M_PI / 180.0;
        const double cx = std::cos(lat0) * std::cos(lon0), cy = std::cos(lat0) * std::sin(lon0), cz = std::sin(lat0);
        auto dens = [&](double a, double b, double c) {
            double d = std::acos(std::max(-1.0, std::min(1.0, a * cx + b * cy + c * cz)));
            return 1.0 + 15.0 * std::exp(-d * d / (2.0 * sig * sig));
        };
        // proposals in blocks; acceptance flags are per proposal, so the accepted order is fixed
        int64_t got = 0, base = 0;
        const int64_t B = 1 << 22;
        std::vector<unsigned char> acc(B);
        while (got < n) {
            parallel_for(B, [&](int64_t k) {
                uint64_t i = (uint64_t)(base + k);
                double u = 2.0 * uni(seed, 1, i) - 1.0, ph = 2.0 * M_PI * uni(seed, 2, i);
                double r = std::sqrt(std::max(0.0, 1.0 - u * u));
                double a = r * std::cos(ph), b = r * std::sin(ph);
                acc[k] = uni(seed, 3, i) * 16.0 < dens(a, b, u);
            });
            for (int64_t k = 0; k < B && got < n; ++k) {
                if (!acc[k]) continue;
                uint64_t i = (uint64_t)(base + k);
                double u = 2.0 * uni(seed, 1, i) - 1.0, ph = 2.0 * M_PI * uni(seed, 2, i);
                double r = std::sqrt(std::max(0.0, 1.0 - u * u));
                x[got] = r * std::cos(ph);
                y[got] = r * std::sin(ph);
                z[got] = u;
                ++got;
            }
            base += B;
        }
        std::vector<double> w(n);
        parallel_for(n, [&](int64_t i) { w[i] = 1.0 / dens(x[i], y[i], z[i]); });
        double tot = 0.0;
        for (int64_t i = 0; i < n; ++i) tot += w[i];
        for (int64_t i = 0; i < n; ++i) area[i] = w[i] * (total_area / tot);
    });
}

int sal_gen_mom6(int nlon, int nlat, double lat0, double lat1, uint64_t seed, double land_frac, double radius,
                 int64_t* n_out, double* x, double* y, double* z, double* area) {
    return guard([&] {
        SAL_REQUIRE(nlon >= 1 && nlat >= 1 && lat1 > lat0 && n_out, SAL_ERR_ARG, "bad arguments");
        SAL_REQUIRE(land_frac >= 0.0 && land_frac < 1.0, SAL_ERR_ARG, "land fraction must be in [0, 1)");
        const int64_t nc = (int64_t)nlon * nlat;
        const double dlon = 2.0 * M_PI / nlon, dlat = (lat1 - lat0) * M_PI / 180.0 / nlat, la0 = lat0 * M_PI / 180.0;
        // smooth seeded field, harmonics l <= 32, amplitude 1/(l+1)
        const int L = 32, nh = (L + 1) * (L + 1);
        std::vector<double> coef(nh);
        for (int l = 0; l <= L; ++l)
            for (int m = -l; m <= l; ++m) coef[l * l + l + m] = gauss(seed, 11, (uint64_t)(l * l + l + m)) / (l + 1.0);
        std::vector<double> field(nc);
        parallel_for(nc, [&](int64_t c) {
            int j = (int)(c / nlon), i = (int)(c % nlon);
            double lat = la0 + (j + 0.5) * dlat, lon = (i + 0.5) * dlon;
            double Y[(L + 1) * (L + 1)];
            real_sh(L, std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), std::sin(lat), Y);
            double f = 0.0;
            for (int k = 0; k < nh; ++k) f += coef[k] * Y[k];
            field[c] = f;
        });
        std::vector<double> sorted(field);
        int64_t qi = std::min<int64_t>(nc - 1, (int64_t)std::floor((1.0 - land_frac) * nc));
        std::nth_element(sorted.begin(), sorted.begin() + qi, sorted.end());
        double thr = sorted[qi];
        int64_t k = 0;
        for (int64_t c = 0; c < nc; ++c) {
            if (field[c] >= thr && land_frac > 0) continue;  // land
            if (x) {
                int j = (int)(c / nlon), i = (int)(c % nlon);
                double lat = la0 + (j + 0.5) * dlat, lon = (i + 0.5) * dlon;
                double s1 = std::sin(la0 + j * dlat), s2 = std::sin(la0 + (j + 1) * dlat);
                x[k] = std::cos(lat) * std::cos(lon);
                y[k] = std::cos(lat) * std::sin(lon);
                z[k] = std::sin(lat);
                area[k] = radius * radius * dlon * (s2 - s1);
            }
            ++k;
        }
        *n_out = k;
    });
}

int sal_gen_ssh(int64_t n, const double* x, const double* y, const double* z, int lmax, double amp, double noise,
                uint64_t seed0, int R, double* ssh) {
    return guard([&] {
        SAL_REQUIRE(n >= 1 && x && y && z && ssh && lmax >= 0 && lmax <= 256 && R >= 1, SAL_ERR_ARG, "bad arguments");
        const int nh = (lmax + 1) * (lmax + 1);
        std::vector<double> coef((size_t)nh * R);  // [k][r]
        for (int r = 0; r < R; ++r)
            for (int l = 0; l <= lmax; ++l)
                for (int m = -l; m <= l; ++m) {
                    int k = l * l + l + m;
                    coef[(size_t)k * R + r] = gauss(seed0 + r, 21, (uint64_t)k) * (amp / (l + 1.0));
                }
        parallel_for(n, [&](int64_t i) {
            std::vector<double> Y(nh), acc(R, 0.0);
            real_sh(lmax, x[i], y[i], z[i], Y.data());
            for (int k = 0; k < nh; ++k) {
                const double yk = Y[k];
                const double* c = &coef[(size_t)k * R];
                for (int r = 0; r < R; ++r) acc[r] += c[r] * yk;
            }
            for (int r = 0; r < R; ++r)
                ssh[(size_t)r * n + i] = acc[r] + (noise > 0 ? noise * gauss(seed0 + r, 22, (uint64_t)i) : 0.0);
        });
    });
}

}  // extern "C"

// ------------------------------------------------------------------------------ kernel timing
#include <atomic>
#include <map>
#include <mutex>

namespace sal {
namespace {
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::atomic<bool> g_prof{false};
thread_local bool t_capturing = false;
std::vector<ProfRec> g_pending;
std::map<std::string, std::pair<double, int64_t>> g_totals;

void prof_collect() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& r : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& t = g_totals[r.name];
            t.first += ms;
            t.second += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_pending.clear();
}
}  // namespace

bool prof_on() { return g_prof && !t_capturing; }
void set_capturing(bool on) { t_capturing = on; }

ProfScope::ProfScope(const char* n, hipStream_t st) : name(n), s(st) {
    if (!prof_on()) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
    (void)hipEventRecord(a, s);
}
ProfScope::~ProfScope() {
    if (!a) return;
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_pending.push_back({name, a, b});
}
}  // namespace sal

extern "C" {

int sal_prof_enable(int on) {
    return guard([&] {
        prof_collect();
        g_prof = on != 0;
    });
}

int sal_prof_reset(void) {
    return guard([&] {
        prof_collect();
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_totals.clear();
    });
}

int sal_prof_count(int* n) {
    return guard([&] {
        SAL_REQUIRE(n, SAL_ERR_ARG, "null pointer");
        prof_collect();
        std::lock_guard<std::mutex> lk(g_prof_mu);
        *n = (int)g_totals.size();
    });
}

int sal_prof_get(int idx, char* name, int len, double* total_ms, int64_t* launches) {
    return guard([&] {
        prof_collect();
        std::lock_guard<std::mutex> lk(g_prof_mu);
        SAL_REQUIRE(idx >= 0 && idx < (int)g_totals.size() && name && len > 0, SAL_ERR_ARG, "bad index");
        auto it = g_totals.begin();
        std::advance(it, idx);
        std::snprintf(name, (size_t)len, "%s", it->first.c_str());
        if (total_ms) *total_ms = it->second.first;
        if (launches) *launches = it->second.second;
    });
}

}  // extern "C"
