// test_ragged_table.cpp -- a C++ host of the ragged solve that takes its
// bucket table from an autotune file through the C ABI, as a C++ caller of
// the library does (no Python involved).
//
//   test_ragged_table TABLE.json SIZES.txt [--fp64]
//
// Loads the table with trid_bucket_table_load, plans the ragged batch whose
// system sizes SIZES.txt lists (whitespace-separated), and prints the plan
// (trid_ragged_plan_info) as one JSON line. With --fp64 it also solves a
// seeded batch of those sizes in the one-pass FP64 mode and prints the
// largest relative residual (trid_residual_f64). Exits non-zero on any
// failed call.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "trid.h"

static int fail(const char* what, int st)
{
    std::fprintf(stderr, "%s: %s\n", what, trid_status_string(st));
    return 1;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s TABLE.json SIZES.txt [--fp64]\n", argv[0]);
        return 2;
    }
    const bool solve = argc > 3 && std::strcmp(argv[3], "--fp64") == 0;
    trid_bucket_cfg table[32];
    int count = 0;
    int st = trid_bucket_table_load(argv[1], table, 32, &count);
    if (st != TRID_OK) return fail("trid_bucket_table_load", st);
    std::vector<int64_t> offs(1, 0);
    std::FILE* f = std::fopen(argv[2], "r");
    if (!f) return fail("open sizes", TRID_EINVAL);
    long long n = 0;
    while (std::fscanf(f, "%lld", &n) == 1) offs.push_back(offs.back() + n);
    std::fclose(f);
    const int64_t N = (int64_t)offs.size() - 1;
    trid_ragged_plan plan = nullptr;
    st = trid_ragged_plan_create(&plan, offs.data(), N, table, count < 32 ? count : 32);
    if (st != TRID_OK) return fail("trid_ragged_plan_create", st);
    trid_ragged_info info{};
    trid_bucket_cfg bk[16];
    int64_t sys[16];
    st = trid_ragged_plan_info(plan, &info, bk, sys, 16);
    if (st != TRID_OK) return fail("trid_ragged_plan_info", st);
    std::printf("{\"table_entries\": %d, \"num_systems\": %lld, \"total_unknowns\": %lld, "
                "\"num_buckets\": %d, \"workgroup_items\": %d, \"wave_items\": %d, "
                "\"grid\": %d, \"buckets\": [",
                count, (long long)info.num_systems, (long long)info.total_unknowns,
                info.num_buckets, info.workgroup_items, info.wave_items, info.grid);
    for (int i = 0; i < info.num_buckets && i < 16; ++i)
        std::printf("%s{\"size\": %d, \"per_lane\": %d, \"lanes\": %d, \"rest\": %d, "
                    "\"systems\": %lld}",
                    i ? ", " : "", bk[i].size, bk[i].per_lane, bk[i].lanes, bk[i].rest,
                    (long long)sys[i]);
    std::printf("]");
    int rc = 0;
    if (solve) {
        const int64_t total = offs.back();
        double* d[5] = {};
        double* relres = nullptr;
        int64_t* d_offs = nullptr;
        for (auto& p : d)
            if (hipMalloc(reinterpret_cast<void**>(&p), total * sizeof(double)) != hipSuccess)
                return fail("hipMalloc", TRID_ENOMEM);
        if (hipMalloc(reinterpret_cast<void**>(&relres), N * sizeof(double)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&d_offs), (N + 1) * sizeof(int64_t)) !=
                hipSuccess)
            return fail("hipMalloc", TRID_ENOMEM);
        (void)hipMemcpy(d_offs, offs.data(), (N + 1) * sizeof(int64_t), hipMemcpyHostToDevice);
        st = trid_generate_f64(0, 5, d[0], d[1], d[2], d[3], total, 1, nullptr);
        if (st == TRID_OK)
            st = trid_ragged_solve_f64(plan, d[0], d[1], d[2], d[3], d[4], 0, TRID_RAGGED_FP64,
                                       nullptr);
        if (st == TRID_OK)
            st = trid_residual_f64(d[0], d[1], d[2], d[3], d[4], d_offs, N, relres, nullptr);
        if (st != TRID_OK) return fail("ragged FP64 solve", st);
        std::vector<double> rr(N);
        (void)hipMemcpy(rr.data(), relres, N * sizeof(double), hipMemcpyDeviceToHost);
        double worst = 0;
        for (double v : rr) worst = v > worst ? v : worst;
        std::printf(", \"max_rel_residual\": %.3e", worst);
        rc = worst <= 1e-12 ? 0 : 1;
        for (auto p : d) (void)hipFree(p);
        (void)hipFree(relres);
        (void)hipFree(d_offs);
    }
    std::printf("}\n");
    trid_ragged_plan_destroy(plan);
    return rc;
}
