// ajbench.cpp -- benchmark driver: builds R and S, selects the join by name and
// runs it, like the reference's `hashing` binary (main -> parse_args ->
// benchmark(), hashing/main.cpp:256-296, 444-705; hashing/benchmark.cpp:164-257).
//
//   ajbench -a PRO|NPO|SHJ_st|SHJ_JM_NP|SHJ_JBCR_NP [-r |R|] [-s |S|] [-x r_seed] [-y s_seed]
//           [-d key distribution: 0 unique/uniform, 2 zipf] [-z zipf theta] [--fk-zipf]
//           [-w window -e step -l interval] [-n nthreads] [-m] [--repeats n]
//           [--pairs] [--out FILE] [-R fileR -J keyR -L tsR] [-S fileS -K keyS -M tsS]
//           [--timing EXP_DIR [--exp-id n] [--gap n]]
//   ajbench -a SERVE [--load f] [--peak Mtps] [--seconds s] [--streams n]
//           [--budget bytes] [--slots n] [--batch-windows n] [--clients n]
//           [--verify n] [--median n] [--sigma s] [--cap-tuples n]
//           [--domain n] [-z theta]      (served windows: ajbench_serve.cpp)
//           [--calibrate n]  (instead of a load point: plain one-window launches
//                             and the server's batched peak on n resident windows)
//
// Sizes: with -w/-e/-l, |R| = |S| = (window / interval) * step as in the
// reference (benchmark.cpp:179-180); otherwise -r/-s (default 128M as in the
// reference's help text, main.cpp:403-404).  -d 0: R keys are a seeded
// permutation of 1..|R|, S keys uniform on 1..|R|.  -d 2, as in the reference
// (createRelation -> create_relation_zipf(rel, n, n, skew), benchmark.cpp:103,
// reached for R at :188 and for S at :204): the keys of each relation are drawn
// Zipf(theta) over 1..its own size.  --fk-zipf keeps R a permutation and draws
// S Zipf(theta) over R's keys (a skewed foreign-key join).  Prints "Results = N"
// like benchmark.cpp:249, plus the checksum and
// the device and end-to-end times.  -m materialises the reference's result
// records (<S.key, S.payloadID> per match in chained per-thread buffers, as a
// JOIN_RESULT_MATERIALIZE build does); --out FILE also writes them as the
// reference's Out.tbl (benchmark.cpp:262-265); --pairs returns the
// (R.payloadID, S.payloadID) pair list instead.  -R/-S load R or S from the
// reference's text files (main.cpp:612-628; key column -J/-K, timestamp column
// -L/-M, |R| and |S| lines from -r/-s) through aj_load_relation, with the
// reference's placement over -n partitions (benchmark.cpp:183-205).
// --timing writes the reference's progressive-result files under EXP_DIR
// (records, timestamps, latency, gaps, breakdown; aj_write_timing_files) with
// lastTS = -w (the window, ms) and prints the progress figures the reference
// prints (t_timer.cpp:67-89).
#include <getopt.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/alliancedb/aj.h"
#include "ajbench_serve.h"



namespace {

struct algo_t {
    const char *name;
    result_t *(*joinAlgo)(relation_t *, relation_t *, int);
};

// the algorithms of this library, selected by name (hashing/main.cpp:306-338)
const algo_t algos[] = {
    {"PRO", PRO},
    {"NPO", NPO},
    {"SHJ_st", SHJ_st},
    {"SHJ_JM_NP", SHJ_JM_NP},
    {"SHJ_JBCR_NP", SHJ_JBCR_NP},
    {"PRO_CPU", PRO_CPU},  // the multithreaded CPU baseline (host threads: -n)
    {"NPO_CPU", NPO_CPU},
    {nullptr, nullptr},
};

struct params {
    const algo_t *algo = &algos[0];
    uint64_t r_size = 128000000, s_size = 128000000;
    uint64_t r_seed = 12345, s_seed = 54321;  // main.cpp:351-352
    int key_distribution = 0;
    double skew = 0.0;
    long window = 0, step = 0, interval = 0;
    int nthreads = 1;
    int materialize = 0;
    int repeats = 1;
    const char *out_file = nullptr;
    const char *file_r = nullptr, *file_s = nullptr;
    int rkey = 0, skey = 0, rts = 0, sts = 0;
    const char *timing_dir = nullptr;
    int exp_id = 0, gap = 1;
    bool fk_zipf = false;  // R unique, S Zipf over R's keys
};

void usage(const char *prog) {
    printf("Usage: %s -a <PRO|NPO|SHJ_st|SHJ_JM_NP|SHJ_JBCR_NP|PRO_CPU|NPO_CPU> [-r R] [-s S] [-x seed] [-y seed]\n"
           "          [-d 0|2] [-z theta] [--fk-zipf]\n"
           "          [-w window -e step -l interval] [-n nthreads] [-m] [--repeats n]\n"
           "          [--pairs] [--out FILE] [-R fileR -J keyR -L tsR] [-S fileS -K keyS -M tsS]\n"
           "          [--timing EXP_DIR [--exp-id n] [--gap n]]\n"
           "       %s -a SERVE [--load f] [--peak Mtps] [--seconds s] [--streams n]\n"
           "          [--budget bytes] [--slots n] [--batch-windows n] [--clients n]\n"
           "          [--verify n] [--median n] [--sigma s] [--cap-tuples n] [--domain n]\n"
           "          [--batch-tuples n] [--calibrate n]\n",
           prog, prog);
}

enum { OPT_LOAD = 1000, OPT_PEAK, OPT_SECONDS, OPT_STREAMS, OPT_BUDGET, OPT_SLOTS, OPT_BWIN,
       OPT_CLIENTS, OPT_VERIFY, OPT_MEDIAN, OPT_SIGMA, OPT_CAPT, OPT_DOMAIN, OPT_BTUP, OPT_CALIB,
       OPT_PAIRS, OPT_OUT, OPT_REPEATS, OPT_TIMING, OPT_EXPID, OPT_GAP, OPT_FKZIPF };

}  // namespace

int main(int argc, char **argv) {
    params p;
    ServeParams sp;
    bool serve = false, skew_given = false;
    static option longopts[] = {
                                {"load", required_argument, 0, OPT_LOAD},
                                {"peak", required_argument, 0, OPT_PEAK},
                                {"seconds", required_argument, 0, OPT_SECONDS},
                                {"streams", required_argument, 0, OPT_STREAMS},
                                {"budget", required_argument, 0, OPT_BUDGET},
                                {"slots", required_argument, 0, OPT_SLOTS},
                                {"batch-windows", required_argument, 0, OPT_BWIN},
                                {"clients", required_argument, 0, OPT_CLIENTS},
                                {"verify", required_argument, 0, OPT_VERIFY},
                                {"median", required_argument, 0, OPT_MEDIAN},
                                {"sigma", required_argument, 0, OPT_SIGMA},
                                {"cap-tuples", required_argument, 0, OPT_CAPT},
                                {"domain", required_argument, 0, OPT_DOMAIN},
                                {"batch-tuples", required_argument, 0, OPT_BTUP},
                                {"calibrate", required_argument, 0, OPT_CALIB},
                                {"pairs", no_argument, 0, OPT_PAIRS},
                                {"out", required_argument, 0, OPT_OUT},
                                {"algo", required_argument, 0, 'a'},
                                {"r-size", required_argument, 0, 'r'},
                                {"s-size", required_argument, 0, 's'},
                                {"r-seed", required_argument, 0, 'x'},
                                {"s-seed", required_argument, 0, 'y'},
                                {"key_distribution", required_argument, 0, 'd'},
                                {"skew", required_argument, 0, 'z'},
                                {"window-size", required_argument, 0, 'w'},
                                {"step-size", required_argument, 0, 'e'},
                                {"interval", required_argument, 0, 'l'},
                                {"nthreads", required_argument, 0, 'n'},
                                {"materialize", no_argument, 0, 'm'},
                                {"repeats", required_argument, 0, OPT_REPEATS},
                                {"timing", required_argument, 0, OPT_TIMING},
                                {"exp-id", required_argument, 0, OPT_EXPID},
                                {"gap", required_argument, 0, OPT_GAP},
                                {"fk-zipf", no_argument, 0, OPT_FKZIPF},
                                {"loadfileR", required_argument, 0, 'R'},
                                {"loadfileS", required_argument, 0, 'S'},
                                {"rkey", required_argument, 0, 'J'},
                                {"skey", required_argument, 0, 'K'},
                                {"rts", required_argument, 0, 'L'},
                                {"sts", required_argument, 0, 'M'},
                                {"help", no_argument, 0, 'h'},
                                {0, 0, 0, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "a:r:s:x:y:d:z:w:e:l:n:mR:S:J:K:L:M:h", longopts, nullptr)) != -1) {
        switch (c) {
            case 'a': {
                if (strcmp(optarg, "SERVE") == 0) {
                    serve = true;
                    break;
                }
                const algo_t *a = algos;
                while (a->name && strcmp(a->name, optarg) != 0) ++a;
                if (!a->name) {
                    fprintf(stderr, "[ERROR] Join algorithm named `%s' does not exist!\n", optarg);
                    usage(argv[0]);
                    return EXIT_FAILURE;
                }
                p.algo = a;
                break;
            }
            case 'r': p.r_size = strtoull(optarg, nullptr, 10); break;
            case 's': p.s_size = strtoull(optarg, nullptr, 10); break;
            case 'x': p.r_seed = strtoull(optarg, nullptr, 10); break;
            case 'y': p.s_seed = strtoull(optarg, nullptr, 10); break;
            case 'd': p.key_distribution = atoi(optarg); break;
            case 'z': p.skew = atof(optarg); skew_given = true; break;
            case OPT_LOAD: sp.load = atof(optarg); break;
            case OPT_PEAK: sp.peak_mtps = atof(optarg); break;
            case OPT_SECONDS: sp.seconds = atof(optarg); break;
            case OPT_STREAMS: sp.streams = atoi(optarg); break;
            case OPT_BUDGET: sp.budget = atof(optarg); break;
            case OPT_SLOTS: sp.slots = atoi(optarg); break;
            case OPT_BWIN: sp.batch_windows = atoi(optarg); break;
            case OPT_CLIENTS: sp.clients = atoi(optarg); break;
            case OPT_VERIFY: sp.verify = atoi(optarg); break;
            case OPT_MEDIAN: sp.median = atof(optarg); break;
            case OPT_SIGMA: sp.sigma = atof(optarg); break;
            case OPT_CAPT: sp.cap_tuples = strtoull(optarg, nullptr, 10); break;
            case OPT_DOMAIN: sp.domain = (uint32_t)strtoul(optarg, nullptr, 10); break;
            case OPT_BTUP: sp.batch_tuples = strtoull(optarg, nullptr, 10); break;
            case OPT_CALIB: sp.calibrate = atoi(optarg); break;
            case 'w': p.window = atol(optarg); break;
            case 'e': p.step = atol(optarg); break;
            case 'l': p.interval = atol(optarg); break;
            case 'n': p.nthreads = atoi(optarg); break;
            case 'm': p.materialize = 2; break;
            case OPT_PAIRS: p.materialize = 1; break;
            case OPT_OUT: p.out_file = optarg; p.materialize = 2; break;
            case OPT_REPEATS: p.repeats = atoi(optarg); break;
            case OPT_TIMING: p.timing_dir = optarg; break;
            case OPT_EXPID: p.exp_id = atoi(optarg); break;
            case OPT_GAP: p.gap = atoi(optarg); break;
            case OPT_FKZIPF: p.fk_zipf = true; break;
            case 'R': p.file_r = optarg; break;
            case 'S': p.file_s = optarg; break;
            case 'J': p.rkey = atoi(optarg); break;
            case 'K': p.skey = atoi(optarg); break;
            case 'L': p.rts = atoi(optarg); break;
            case 'M': p.sts = atoi(optarg); break;
            case 'h': usage(argv[0]); return EXIT_SUCCESS;
            default: usage(argv[0]); return EXIT_FAILURE;
        }
    }
    if (serve) {
        if (skew_given) sp.theta = p.skew;
        if (sp.streams < 1 || sp.clients < 1 || !(sp.load > 0) || !(sp.seconds > 0) ||
            sp.cap_tuples < 1 || sp.domain < 1) {
            fprintf(stderr, "[ERROR] bad SERVE parameters\n");
            return EXIT_FAILURE;
        }
        return run_serve(sp);
    }
    if (p.window > 0 && p.step > 0 && p.interval > 0) {
        p.r_size = p.s_size = (uint64_t)(p.window / p.interval) * (uint64_t)p.step;
    }
    if (p.key_distribution != 0 && p.key_distribution != 2) {
        fprintf(stderr, "[ERROR] key distribution must be 0 (unique/uniform) or 2 (zipf)\n");
        return EXIT_FAILURE;
    }
    printf("[INFO ] %s relations R=%llu S=%llu (seeds %llu/%llu)\n",
           p.file_r || p.file_s ? "Loading" : "Creating", (unsigned long long)p.r_size,
           (unsigned long long)p.s_size, (unsigned long long)p.r_seed,
           (unsigned long long)p.s_seed);
    std::vector<tuple_t> R(p.file_r ? 0 : p.r_size), S(p.file_s ? 0 : p.s_size);
    relation_t relR = {R.data(), p.r_size, nullptr};
    relation_t relS = {S.data(), p.s_size, nullptr};
    relation_payload_t plR = {nullptr, 0}, plS = {nullptr, 0};
    const uint32_t parts = (uint32_t)(p.nthreads > 0 ? p.nthreads : 1);
    int rc = AJ_OK;
    const bool zipf_both = p.key_distribution == 2 && !p.fk_zipf;
    if (p.file_r)
        rc = aj_load_relation(&relR, &plR, p.rkey, p.rts, p.file_r, p.r_size, parts);
    else if (zipf_both)
        rc = aj_gen_zipf(R.data(), p.r_size, (uint32_t)(p.r_size ? p.r_size : 1), p.skew, p.r_seed);
    else
        rc = aj_gen_unique(R.data(), p.r_size, p.r_seed);
    if (rc == AJ_OK && p.file_s) {
        rc = aj_load_relation(&relS, &plS, p.skey, p.sts, p.file_s, p.s_size, parts);
    } else if (rc == AJ_OK) {
        if (zipf_both)
            rc = aj_gen_zipf(S.data(), p.s_size, (uint32_t)(p.s_size ? p.s_size : 1), p.skew, p.s_seed);
        else if (p.key_distribution == 2)
            rc = aj_gen_zipf(S.data(), p.s_size, (uint32_t)(p.r_size ? p.r_size : 1), p.skew, p.s_seed);
        else
            rc = aj_gen_uniform(S.data(), p.s_size, 1, (uint32_t)(p.r_size ? p.r_size : 1), p.s_seed);
    }
    if (rc != AJ_OK) {
        fprintf(stderr, "[ERROR] input relations: %s\n", aj_last_error());
        return EXIT_FAILURE;
    }
    aj_set_option("materialize", p.materialize);
    aj_set_option("timing", p.timing_dir ? 1 : 0);
    printf("[INFO ] Running join algorithm %s ...\n", p.algo->name);
    for (int rep = 0; rep < p.repeats; ++rep) {
        auto t0 = std::chrono::steady_clock::now();
        result_t *res = p.algo->joinAlgo(&relR, &relS, p.nthreads);
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!res) {
            fprintf(stderr, "[ERROR] %s failed: %s\n", p.algo->name, aj_last_error());
            return EXIT_FAILURE;
        }
        const aj_result_t *x = aj_result_ext(res);
        printf("[INFO ] Results = %lld. checksum = %llu pairs = %llu device_ms = %.3f "
               "total_ms = %.3f wall_ms = %.3f\n",
               (long long)res->totalresults, (unsigned long long)x->checksum,
               (unsigned long long)x->npairs, x->device_ms, x->total_ms, ms);
        if (p.timing_dir && rep + 1 == p.repeats) {
            // the lazy joins' result names (PRJ / NPJ); the eager joins use their
            // algorithm name, as the reference's finishing does (param.algo_name)
            const char *tname = strcmp(p.algo->name, "PRO") == 0   ? "PRJ"
                                : strcmp(p.algo->name, "NPO") == 0 ? "NPJ"
                                                                   : p.algo->name;
            aj_timing_summary_t sm;
            relR.payload = p.file_r ? &plR : nullptr;
            relS.payload = p.file_s ? &plS : nullptr;
            if (aj_write_timing_files(res, &relR, &relS, p.timing_dir, tname,
                                      p.exp_id, p.window, p.gap, &sm) != AJ_OK) {
                fprintf(stderr, "[ERROR] %s\n", aj_last_error());
                aj_free_result(res);
                return EXIT_FAILURE;
            }
            printf("Time to obtain 1%%, 5%%, 10%%, 25%%, 50%% of results (MSECS): \n");
            printf("(%.2f) \t (%.2f) \t (%.2f) \t (%.2f) \t (%.2f)\n\n", sm.progress_ms[0],
                   sm.progress_ms[1], sm.progress_ms[2], sm.progress_ms[3], sm.progress_ms[4]);
            printf("95th latency: (%.2f)\t 99th latency: (%.2f)\n", sm.latency95_ms,
                   sm.latency99_ms);
        }
        if (p.out_file && rep + 1 == p.repeats) {
            printf("[INFO ] Persisting the join result to \"%s\" ...\n", p.out_file);
            if (aj_write_result_file(res, p.out_file) != AJ_OK) {
                fprintf(stderr, "[ERROR] %s\n", aj_last_error());
                aj_free_result(res);
                return EXIT_FAILURE;
            }
        }
        aj_free_result(res);
    }
    if (p.file_r) {
        free(relR.tuples);
        free(plR.ts);
    }
    if (p.file_s) {
        free(relS.tuples);
        free(plS.ts);
    }
    printf("[INFO ] DONE.\n");
    return EXIT_SUCCESS;
}
