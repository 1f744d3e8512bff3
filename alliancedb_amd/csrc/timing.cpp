// timing.cpp -- progressive-result files of a timed drop-in result, in the
// layout of the reference's timing build.
//
// The reference samples every record_gap-th match of each joining thread with
// the cycle counter (END_PROGRESSIVE_MEASURE, timer/t_timer.h:195-211; the
// probe side is S: parallel_radix_join.cpp:341-342, common_functions.cpp:272),
// turns the samples into match times, latencies and disorder gaps (merge,
// t_timer.cpp:401-437), sorts each (sortRecords, :444-452) and writes them with
// the record counts (dump_timing, :31-99); breakdown_global (:113-140) writes
// the per-phase cycles per input.  Here the samples come from the timed records
// of the device join (device clock ticks since the join started, per join task
// or probe block) and the phase times from the device events of the counting
// join; they are converted to the reference's 2.1 GHz cycles, and every formula
// after that follows the reference's arithmetic (unsigned 64-bit cycle
// differences, double divisions by 2.1e6, std::to_string of doubles).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "aj_internal.h"

namespace aj {
namespace {

constexpr double kCyclesPerMs = 2.1 * 1E6;

int make_dirs(const std::string &path) {
    for (size_t i = 1; i <= path.size(); ++i) {
        if (i < path.size() && path[i] != '/') continue;
        const std::string sub = path.substr(0, i);
        if (mkdir(sub.c_str(), 0755) && errno != EEXIST)
            return set_error(AJ_ERR_ARG, "cannot create %s", sub.c_str());
    }
    return AJ_OK;
}

// One "%f" line per value (std::to_string of a double), like the reference's
// ofstream << std::to_string(element + lastTS) << "\n".
int write_doubles(const std::string &path, const std::vector<double> &v, long add) {
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return set_error(AJ_ERR_ARG, "cannot write %s", path.c_str());
    std::vector<char> buf(1 << 20);
    size_t used = 0;
    bool bad = false;
    for (double x : v) {
        if (used + 400 > buf.size()) {
            bad |= fwrite(buf.data(), 1, used, f) != used;
            used = 0;
        }
        const int k = snprintf(buf.data() + used, buf.size() - used, "%f\n", x + add);
        used += (size_t)std::max(k, 0);
    }
    if (used) bad |= fwrite(buf.data(), 1, used, f) != used;
    bad |= ferror(f) != 0;
    fclose(f);
    return bad ? set_error(AJ_ERR_ARG, "error writing %s", path.c_str()) : AJ_OK;
}

}  // namespace
}  // namespace aj

using namespace aj;

extern "C" int aj_write_timing_files(const result_t *r, const relation_t *relR,
                                     const relation_t *relS, const char *exp_dir,
                                     const char *name, int exp_id, long lastTS, int gap,
                                     aj_timing_summary_t *summary) {
    if (!r || !relR || !relS || !exp_dir || !name || gap < 1)
        return set_error(AJ_ERR_ARG, "aj_write_timing_files: bad arguments");
    const aj_result_t *x = reinterpret_cast<const aj_result_t *>(r);
    if (r->totalresults > 0 && (x->ntimed != (uint64_t)r->totalresults || !x->timed))
        return set_error(AJ_ERR_STATE, "result carries no timed records (option timing 1)");
    if (r->totalresults > 0 && !(x->clock_hz > 0))
        return set_error(AJ_ERR_STATE, "result carries no device clock rate");
    if (x->ntimed_r > x->ntimed) return set_error(AJ_ERR_STATE, "bad count of R-side records");
    const uint64_t *sts = relS->payload ? relS->payload->ts : nullptr;
    const uint64_t *rts = relR->payload ? relR->payload->ts : nullptr;

    // production order: device time, then side (S before R), then payloadID;
    // every gap-th sampled.  The last ntimed_r records are R tuples' probes
    // (SHJ); merge treats the R and S samples alike but against their own
    // side's timestamps and with their own sample index (t_timer.cpp:412-434).
    struct Rec {
        uint64_t ticks;
        int r;
        uint32_t id;
        bool operator<(const Rec &o) const {
            return ticks != o.ticks ? ticks < o.ticks : r != o.r ? r < o.r : id < o.id;
        }
    };
    std::vector<Rec> rec(x->ntimed);
    const uint64_t first_r = x->ntimed - x->ntimed_r;
    for (uint64_t k = 0; k < x->ntimed; ++k)
        rec[k] = {x->timed[k] >> 32, k >= first_r, (uint32_t)x->timed[k]};
    std::sort(rec.begin(), rec.end());
    std::vector<double> times, lat, gaps;
    const uint64_t wait_cycles = (uint64_t)(lastTS * 2.1 * 1E6);
    const double cycles_per_tick = 2.1e9 / x->clock_hz;  // 21 exactly for a 100 MHz clock
    int64_t idx[2] = {0, 0};
    for (uint64_t k = (uint64_t)gap - 1; k < rec.size(); k += (uint64_t)gap) {
        const Rec &e = rec[k];
        const relation_t *rel = e.r ? relR : relS;
        const uint64_t *ts = e.r ? rts : sts;
        if (ts && e.id >= rel->num_tuples)
            return set_error(AJ_ERR_ARG, "%s.payloadID %u outside the relation's %llu timestamps",
                             e.r ? "R" : "S", e.id, (unsigned long long)rel->num_tuples);
        const uint64_t cycles = (uint64_t)((double)e.ticks * cycles_per_tick);
        times.push_back(cycles / (2.1 * 1E6));
        const int64_t latency = (int64_t)(cycles - (ts ? ts[e.id] : 0) + wait_cycles);
        lat.push_back(latency / (2.1 * 1E6));
        gaps.push_back((double)((int32_t)e.id - (int32_t)idx[e.r]++));
    }
    std::sort(times.begin(), times.end());
    std::sort(lat.begin(), lat.end());
    std::sort(gaps.begin(), gaps.end());

    const std::string root = std::string(exp_dir) + "/results/";
    const std::string file = std::string(name) + "_" + std::to_string(exp_id) + ".txt";
    for (const char *d : {"records", "timestamps", "latency", "gaps", "breakdown"})
        AJ_TRY(make_dirs(root + d));
    {
        const unsigned long inputs = (unsigned long)(relR->num_tuples + relS->num_tuples);
        FILE *f = fopen((root + "records/" + file).c_str(), "w");
        if (!f) return set_error(AJ_ERR_ARG, "cannot write %srecords/%s", root.c_str(), file.c_str());
        fprintf(f, "%lu\n%lld\n", inputs, (long long)r->totalresults);
        if (fclose(f)) return set_error(AJ_ERR_ARG, "error writing records file");
    }
    AJ_TRY(write_doubles(root + "timestamps/" + file, times, lastTS));
    AJ_TRY(write_doubles(root + "latency/" + file, lat, lastTS));
    AJ_TRY(write_doubles(root + "gaps/" + file, gaps, lastTS));

    {  // breakdown_global of the lazy joins, t_timer.cpp:113-140
        auto cyc = [](float ms) { return (uint64_t)((double)ms * kCyclesPerMs); };
        const float *ph = x->phase_ms;
        uint64_t wait = 0, part = 0, build = 0, join = 0;
        if (x->algo == AJ_ALGO_NPO) {
            build = cyc(ph[AJ_PH_NPO_BUILD]);
            join = cyc(ph[AJ_PH_NPO_PROBE]);
        } else if (x->algo == AJ_ALGO_SHJ) {
            join = cyc(x->device_ms);  // an eager join builds and probes in one phase
        } else {
            float p = 0;
            for (int i = AJ_PH_PLAN1; i <= AJ_PH_SCATTER2; ++i) p += ph[i];
            part = cyc(p);
            join = cyc(ph[AJ_PH_TASKS] + ph[AJ_PH_JOIN]);
        }
        uint64_t overall = cyc(x->device_ms);
        if (lastTS != 0) {
            wait = (uint64_t)(lastTS * 2.1 * 1E6);
            overall += wait;
        }
        const uint64_t others = overall - (wait + part + build + 0 + 0 + join);
        const int64_t total = (int64_t)(relR->num_tuples + relS->num_tuples);
        FILE *f = fopen((root + "breakdown/" + file).c_str(), "w");
        if (!f) return set_error(AJ_ERR_ARG, "cannot write the breakdown file");
        if (total > 0)
            fprintf(f, "%f\n%f\n%f\n%f\n%f\n%f\n%lu\n", (double)wait / total, (double)part / total,
                    (double)build / total, 0.0, 0.0, (double)join / total,
                    (unsigned long)(others / (uint64_t)total));
        if (fclose(f)) return set_error(AJ_ERR_ARG, "error writing the breakdown file");
    }

    if (summary) {  // the figures dump_timing prints, t_timer.cpp:37-47, 67-89
        memset(summary, 0, sizeof(*summary));
        const int n = (int)times.size() - 1;
        const double pct[5] = {0.01, 0.05, 0.10, 0.25, 0.50};
        for (int i = 0; i < 5; ++i) {
            const int c = (int)ceil(n * pct[i]);
            summary->progress_ms[i] = times.size() > (size_t)(long)c && c >= 0 ? times[c] + lastTS : 0;
        }
        const int c95 = (int)std::min(ceil(n * 0.95), double(n - 1));
        const int c99 = (int)std::min(ceil(n * 0.99), double(n - 1));
        summary->latency95_ms = c95 >= 0 && times.size() > (size_t)c95 ? lat[c95] : 0;
        summary->latency99_ms = c99 >= 0 && times.size() > (size_t)c99 ? lat[c99] : 0;
        summary->samples = times.size();
    }
    return AJ_OK;
}
