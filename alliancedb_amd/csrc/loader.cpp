// loader.cpp -- relations from the reference's on-disk text format.
//
// Restates load_relation (hashing/utils/generator.cpp:645-666), read_relation
// (:899-1027) and smooth (:1050-1067):
//   * the field separator is the first ',' or '|' anywhere in the file
//     (:922-938); a file with neither is an error;
//   * line i (0-based, no header skipped) gives the key, std::stoi of field
//     `keyby` (:984), and, when tsKey != 0, the timestamp
//     stol(field tsKey) * 2.1 * 1E6 (:987-988), evaluated left to right in
//     doubles and truncated to uint64;
//   * the timestamps are smoothed: each run of equal values is spread linearly
//     towards the next distinct value, ts[c + j] += j * (next - ts[c]) / run;
//     the trailing run stays as it is (:1050-1067);
//   * line i goes to partition i % partitions, or to the last partition once
//     that one is full, at the partition's next free index; partitions are
//     n / partitions tuples, the last one takes the remainder (:964-1024);
//     payloadID = the index it lands on, and ts[index] = its timestamp.
// Where the reference's behaviour is undefined this returns an error instead:
// unreadable file, no separator, fewer lines than num_tuples, a missing or
// unparsable field, a key outside int32 (stoi throws), a negative timestamp,
// smoothing that overflows uint64 (a decreasing timestamp after a run of equal
// ones), partitions == 0.  tsKey == 0 leaves every timestamp 0 (the reference
// leaves them uninitialised).
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "aj_internal.h"

namespace aj {
namespace {

bool read_file(const char *path, std::vector<char> *out) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    out->clear();
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out->insert(out->end(), buf, buf + n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

// Field k of [b, e) split at sep (a field may be empty; the last runs to e).
bool field(const char *b, const char *e, char sep, int k, const char **fb, const char **fe) {
    for (int i = 0; i < k; ++i) {
        const char *p = (const char *)memchr(b, sep, (size_t)(e - b));
        if (!p) return false;
        b = p + 1;
    }
    const char *p = (const char *)memchr(b, sep, (size_t)(e - b));
    *fb = b;
    *fe = p ? p : e;
    return true;
}

// strtol on a NUL-terminated copy: std::stoi / std::stol parse the same way
// (leading white space, optional sign, decimal digits, anything after ignored).
bool parse_long(const char *b, const char *e, long *v) {
    std::string tmp(b, e);
    char *end = nullptr;
    errno = 0;
    const long x = strtol(tmp.c_str(), &end, 10);
    if (end == tmp.c_str() || errno == ERANGE) return false;
    *v = x;
    return true;
}

}  // namespace
}  // namespace aj

using namespace aj;

extern "C" int aj_load_relation(relation_t *rel, relation_payload_t *pl, int32_t keyby,
                                int32_t tsKey, const char *filename, uint64_t num_tuples,
                                uint32_t partitions) {
    if (!rel || !filename || keyby < 0 || tsKey < 0)
        return set_error(AJ_ERR_ARG, "aj_load_relation: bad arguments");
    if (partitions == 0) return set_error(AJ_ERR_ARG, "aj_load_relation: partitions must be >= 1");
    if (num_tuples >= (1ull << 31))
        return set_error(AJ_ERR_ARG, "aj_load_relation: %llu tuples (limit 2^31-1)",
                         (unsigned long long)num_tuples);
    std::vector<char> text;
    if (!read_file(filename, &text)) return set_error(AJ_ERR_ARG, "cannot read %s", filename);
    char sep = 0;
    for (char c : text)
        if (c == ',' || c == '|') {
            sep = c;
            break;
        }
    if (!sep) return set_error(AJ_ERR_ARG, "%s: no ',' or '|' separated fields", filename);

    const uint64_t n = num_tuples;
    std::vector<int32_t> key(n);
    std::vector<uint64_t> ts(n, 0);
    const char *p = text.data(), *end = text.data() + text.size();
    uint64_t i = 0;
    for (; i < n && p < end; ++i) {
        const char *nl = (const char *)memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        const char *fb, *fe;
        long v;
        if (!field(p, le, sep, keyby, &fb, &fe) || !parse_long(fb, fe, &v) || v < INT_MIN ||
            v > INT_MAX)
            return set_error(AJ_ERR_ARG, "%s line %llu: no int32 key in field %d", filename,
                             (unsigned long long)i + 1, keyby);
        key[i] = (int32_t)v;
        if (tsKey != 0) {
            if (!field(p, le, sep, tsKey, &fb, &fe) || !parse_long(fb, fe, &v))
                return set_error(AJ_ERR_ARG, "%s line %llu: no timestamp in field %d", filename,
                                 (unsigned long long)i + 1, tsKey);
            if (v < 0)
                return set_error(AJ_ERR_ARG, "%s line %llu: negative timestamp", filename,
                                 (unsigned long long)i + 1);
            const double t = (double)v * 2.1 * 1E6;
            if (t >= 18446744073709551616.0)
                return set_error(AJ_ERR_ARG, "%s line %llu: timestamp out of range", filename,
                                 (unsigned long long)i + 1);
            ts[i] = (uint64_t)t;
        }
        p = nl ? nl + 1 : end;
    }
    if (i < n)
        return set_error(AJ_ERR_ARG, "%s holds %llu lines, %llu tuples requested", filename,
                         (unsigned long long)i, (unsigned long long)n);

    // smooth: spread each run of equal timestamps towards the next value
    if (n) {
        uint64_t run = 0, cur = 0, head = ts[0];
        for (uint64_t k = 0; k < n; ++k) {
            if (ts[k] == head) {
                ++run;
                continue;
            }
            const double inner = (double)(ts[k] - ts[cur]) / (double)run;
            for (uint64_t j = 0; j < run; ++j) {
                // two statements: a product and a sum, each rounded (no fused multiply-add)
                const double add = (double)j * inner;
                const double t = (double)ts[cur + j] + add;
                if (t >= 18446744073709551616.0)
                    return set_error(AJ_ERR_ARG,
                                     "%s: timestamp smoothing overflows at line %llu "
                                     "(timestamps decrease)",
                                     filename, (unsigned long long)(cur + j) + 1);
                ts[cur + j] = (uint64_t)t;
            }
            head = ts[k];
            cur = k;
            run = 1;
        }
    }

    void *tup = nullptr;
    if (posix_memalign(&tup, 64, std::max<uint64_t>(n, 1) * sizeof(tuple_t)))
        return set_error(AJ_ERR_NOMEM, "host allocation of %llu tuples failed",
                         (unsigned long long)n);
    uint64_t *dts = nullptr;
    bool own_ts = false;
    if (pl) {
        dts = pl->ts;
        if (!dts) {
            void *q = nullptr;
            if (posix_memalign(&q, 64, std::max<uint64_t>(n, 1) * sizeof(uint64_t))) {
                free(tup);
                return set_error(AJ_ERR_NOMEM, "host allocation of %llu timestamps failed",
                                 (unsigned long long)n);
            }
            dts = (uint64_t *)q;
            own_ts = true;
        }
    }
    // round-robin placement over the partitions' index ranges
    const uint64_t P = partitions, per = n / P;
    std::vector<uint64_t> fill(P, 0);
    tuple_t *t = (tuple_t *)tup;
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t part = k % P;
        const uint64_t part_end = part == P - 1 ? n : per * (part + 1);
        if (per * part + fill[part] == part_end) part = P - 1;
        const uint64_t idx = per * part + fill[part]++;
        t[idx].key = key[k];
        t[idx].payloadID = (int32_t)idx;
        if (dts) dts[idx] = ts[k];
    }
    rel->tuples = t;
    rel->num_tuples = n;
    if (pl) {
        if (own_ts) pl->ts = dts;
        pl->num_tuples = n;
    }
    return AJ_OK;
}
