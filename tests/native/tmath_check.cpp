// Accuracy of tdt::tm (csrc/tmath.hpp) against glibc libm, in ulps of the glibc result.
// Prints one line per function: name max_ulp mean_ulp samples.
#include <cmath>
#include <cstdio>
#include <random>

#include "tmath.hpp"

static double ulp_err(double got, double want) {
    if (std::isnan(want)) return std::isnan(got) ? 0.0 : 1e300;
    if (std::isinf(want)) return got == want ? 0.0 : 1e300;
    double u = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
    if (want == 0.0) return got == 0.0 ? 0.0 : std::fabs(got) / 4.9e-324;
    return std::fabs(got - want) / u;
}

template <class F, class G>
static void run(const char *name, F f, G g, double lo, double hi, bool logscale, int n) {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    double mx = 0.0, sum = 0.0, worst = 0.0;
    for (int k = 0; k < n; ++k) {
        double x = logscale ? std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * u(rng))
                            : lo + (hi - lo) * u(rng);
        double e = ulp_err(f(x), g(x));
        sum += e;
        if (e > mx) { mx = e; worst = x; }
    }
    std::printf("%s %.3f %.4f %d %.17g\n", name, mx, sum / n, n, worst);
}

int main() {
    const int n = 400000;
    using namespace tdt;
    run("tanh_small", [](double x) { return tm::tanh(x); }, [](double x) { return std::tanh(x); }, 1e-12, 0.5, true, n);
    run("tanh_mid", [](double x) { return tm::tanh(x); }, [](double x) { return std::tanh(x); }, 0.5, 25.0, false, n);
    run("tanh_neg", [](double x) { return tm::tanh(-x); }, [](double x) { return std::tanh(-x); }, 1e-6, 5.0, true, n);
    run("pow_0.6", [](double x) { return tm::pow(x, 0.6); }, [](double x) { return std::pow(x, 0.6); }, 1.0, 5000.0, true, n);
    run("pow_-2.5", [](double x) { return tm::pow(x, -2.5); }, [](double x) { return std::pow(x, -2.5); }, 0.01, 50.0, true, n);
    run("pow_1/7", [](double x) { return tm::pow(x, 1.0 / 7.0); }, [](double x) { return std::pow(x, 1.0 / 7.0); }, 1e-6, 1.0, true, n);
    run("pow_3.5", [](double x) { return tm::pow(x, 3.5); }, [](double x) { return std::pow(x, 3.5); }, 0.5, 4.0, false, n);
    run("exp", [](double x) { return tm::exp(x); }, [](double x) { return std::exp(x); }, -700.0, 700.0, false, n);
    run("log", [](double x) { return tm::log_pos(x); }, [](double x) { return std::log(x); }, 1e-300, 1e300, true, n);
    // special values
    int bad = 0;
    bad += tm::tanh(0.0) != 0.0;
    bad += tm::tanh(1e300) != 1.0;
    bad += !std::isnan(tm::tanh(NAN));
    bad += tm::pow(0.0, 0.6) != 0.0;
    bad += !std::isnan(tm::pow(-1.0, 0.6));
    bad += !std::isnan(tm::pow(NAN, 0.6));
    bad += tm::exp(-1000.0) != 0.0;
    bad += tm::exp(1000.0) != INFINITY;
    std::printf("special %d 0 0 0\n", bad);
    return 0;
}
