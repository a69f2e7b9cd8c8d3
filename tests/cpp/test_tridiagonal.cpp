// test_tridiagonal.cpp -- C++ end-to-end driver written against the
// reference's solver names (trid_compat.h), in the shape of the reference
// driver test_tridiagonal.cpp:28-118: generate systems with the reference's
// value ranges (b in [8,9], a in [3,4], c in [2,3], d in [5,6]), solve on the
// GPU, solve with the CPU serial Thomas, compare per system by RMS <= 1e-3
// with NaN counted as failure (tridiagonal_gold.h:111-148).
//
// Unlike the reference driver (which always returns EXIT_SUCCESS and runs
// only pcrGaussWrapper(..., 64)), every method runs -- including the
// partial-pivoting MKL baseline's names, tridiagonalMKL and
// multiThreadTridiagonalMKL -- and the process
// exits with the number of failed (method, system) pairs, capped at 255.
//
// Test infrastructure: links the CPU oracle (oracle/liboracle.so) as the
// checker. Usage: test_tridiagonal [numSystems] [systemSize] [f32|f64]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "trid_compat.h"

extern "C" {
int oracle_generate_f32(int kind, uint64_t seed, float* a, float* b, float* c, float* d,
                        const int64_t* offsets, int64_t num_systems);
int oracle_generate_f64(int kind, uint64_t seed, double* a, double* b, double* c, double* d,
                        const int64_t* offsets, int64_t num_systems);
int oracle_thomas_f64_from_f32(const float*, const float*, const float*, const float*, double*,
                               int64_t, int64_t, int);
int oracle_thomas_f64(const double*, const double*, const double*, const double*, double*,
                      int64_t, int64_t, int);
int64_t oracle_compare_ragged(const double* x, const double* xref, const int64_t* offsets,
                              int64_t num_systems, double rms_eps, double* relerr, double* rms);
}

static void gen(int kind, uint64_t seed, std::vector<float>* v, const int64_t* off, int64_t N)
{
    oracle_generate_f32(kind, seed, v[0].data(), v[1].data(), v[2].data(), v[3].data(), off, N);
}
static void gen(int kind, uint64_t seed, std::vector<double>* v, const int64_t* off, int64_t N)
{
    oracle_generate_f64(kind, seed, v[0].data(), v[1].data(), v[2].data(), v[3].data(), off, N);
}
static void gold(std::vector<float>* v, double* x, int n, int N)
{
    oracle_thomas_f64_from_f32(v[0].data(), v[1].data(), v[2].data(), v[3].data(), x, n, N, 0);
}
static void gold(std::vector<double>* v, double* x, int n, int N)
{
    oracle_thomas_f64(v[0].data(), v[1].data(), v[2].data(), v[3].data(), x, n, N, 0);
}

template <typename T>
static int run(int numSystems, int systemSize)
{
    const size_t total = (size_t)numSystems * systemSize;
    std::vector<T> arr[4];
    for (auto& v : arr) v.assign(total, T(0));
    std::vector<int64_t> off(numSystems + 1);
    for (int s = 0; s <= numSystems; ++s) off[s] = (int64_t)s * systemSize;
    gen(2, 2010, arr, off.data(), numSystems);  // reference value ranges
    std::vector<double> xref(total);
    gold(arr, xref.data(), systemSize, numSystems);

    int failures = 0;
    T *a = arr[0].data(), *b = arr[1].data(), *c = arr[2].data(), *d = arr[3].data();
    std::vector<T> x(total);
    std::vector<double> xd(total), rel(numSystems), rms(numSystems);
    const int n_pad = 1 << (int)std::ceil(std::log2((double)systemSize));
    struct Case {
        std::string name;
        int param;
    };
    std::vector<Case> cases = {{"cr", 0}, {"pcr", 0}, {"crNBC", 0}, {"crNBC32", 0}};
    for (int rest = 2; rest <= n_pad / 2 && rest <= 256; rest *= 2) {
        cases.push_back({"crpcr", rest});
        cases.push_back({"crpcrNBC", rest});
    }
    for (int g = 1; g <= n_pad / 2 && g <= 512; g *= 2) cases.push_back({"pcrGaussWrapper", g});
    cases.push_back({"tridiagonalMKL", 0});
    cases.push_back({"multiThreadTridiagonalMKL", 0});
    std::vector<T> dwork(total);

    for (const Case& k : cases) {
        std::fill(x.begin(), x.end(), T(NAN));
        const T* xs = x.data();
        if (k.name == "cr") cr(a, b, c, d, x.data(), systemSize, numSystems);
        else if (k.name == "pcr") pcr(a, b, c, d, x.data(), systemSize, numSystems);
        else if (k.name == "crNBC") crNBC(a, b, c, d, x.data(), systemSize, numSystems);
        else if (k.name == "crNBC32") crNBC32(a, b, c, d, x.data(), systemSize, numSystems);
        else if (k.name == "crpcr") crpcr(a, b, c, d, x.data(), systemSize, numSystems, k.param);
        else if (k.name == "crpcrNBC")
            crpcrNBC(a, b, c, d, x.data(), systemSize, numSystems, k.param);
        else if (k.name == "pcrGaussWrapper")
            pcrGaussWrapper(a, b, c, d, x.data(), systemSize, numSystems, k.param);
        else {
            // solves in the d buffer and swaps the pointers (tridiagonalMKL.cpp:22-25,
            // 56-59)
            dwork = arr[3];
            T* dp = dwork.data();
            T* xp = x.data();
            if (k.name == "tridiagonalMKL")
                tridiagonalMKL(a, b, c, dp, xp, systemSize, numSystems);
            else
                multiThreadTridiagonalMKL(a, b, c, dp, xp, systemSize, numSystems);
            if (trid_last_status() == TRID_OK && (xp != dwork.data() || dp != x.data())) {
                std::printf("%s did not swap d and x\n", k.name.c_str());
                failures += numSystems;
                continue;
            }
            xs = xp;
        }
        const int st = trid_last_status();
        if (st == TRID_ENOTSUP) {
            std::printf("%-16s %5d  not supported at n=%d\n", k.name.c_str(), k.param,
                        systemSize);
            continue;
        }
        if (st != TRID_OK) {
            std::printf("%-16s %5d  status %s\n", k.name.c_str(), k.param,
                        trid_status_string(st));
            failures += numSystems;
            continue;
        }
        for (size_t i = 0; i < total; ++i) xd[i] = (double)xs[i];
        const int64_t f = oracle_compare_ragged(xd.data(), xref.data(), off.data(), numSystems,
                                                1e-3, rel.data(), rms.data());
        double worst = 0.0;
        for (int s = 0; s < numSystems; ++s)
            if (!(rel[s] <= worst)) worst = rel[s];
        std::printf("%-16s %5d  max rel err %.3e  rms failures %lld\n", k.name.c_str(), k.param,
                    worst, (long long)f);
        failures += (int)f;
    }
    return failures;
}

int main(int argc, char** argv)
{
    const int numSystems = argc > 1 ? std::atoi(argv[1]) : 2;
    const int systemSize = argc > 2 ? std::atoi(argv[2]) : 8192;
    const bool dbl = argc > 3 && std::strcmp(argv[3], "f64") == 0;
    if (trid_device_count() < 1) {
        std::printf("no gfx950 device\n");
        return 2;
    }
    std::printf("numSystems %d systemSize %d %s\n", numSystems, systemSize, dbl ? "f64" : "f32");
    const int failures = dbl ? run<double>(numSystems, systemSize) : run<float>(numSystems, systemSize);
    std::printf("%d failures\n", failures);
    return failures > 255 ? 255 : failures;
}
