// ajbench_serve.h -- parameters of the served window-join load generator
// (ajbench -a SERVE, ajbench_serve.cpp).
#pragma once

#include <cstdint>

struct ServeParams {
    int streams = 64;
    double load = 0.5;           // fraction of peak
    double peak_mtps = 55000.0;  // M tuples/s (R+S) the load is relative to
    double seconds = 2.0;
    double median = 262144, sigma = 1.0;
    uint64_t cap_tuples = 1ull << 24;
    uint32_t domain = 1u << 24;
    double theta = 1.0;
    double budget = 200e9;
    int slots = 3;
    int batch_windows = 256;
    uint64_t batch_tuples = 1ull << 27;
    int clients = 4;
    int verify = 64;
    uint64_t seed = 2024;
    // > 0: instead of a load point, measure the two reference rates on this many
    // windows of the same distribution resident in HBM -- plain one-window
    // launches back to back, and the server's batched peak
    int calibrate = 0;
};

int run_serve(const ServeParams &P);
