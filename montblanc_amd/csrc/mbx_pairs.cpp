// Host-side pair-block planner (no device code).
//
// Groups the baselines of one time step into blocks of pb x qb antenna
// pairs keyed by (p / pb, q / qb); a pair that lands on an occupied slot
// (a duplicated baseline) opens another block with the same key. Blocks
// are emitted in key order, so neighbouring threads share antennas.
#include <map>
#include <tuple>
#include <vector>

#include "montblanc_amd.h"

namespace mbx {
int set_error(int code, const char *fmt, ...);
}

extern "C" {

int64_t mbx_pair_plan_capacity(int32_t nbl, int32_t pb, int32_t qb) {
    if (nbl < 0 || pb <= 0 || qb <= 0) return 0;
    return int64_t(nbl) * (2 + int64_t(pb) * qb);
}

int64_t mbx_build_pair_plan(const int32_t *a1, const int32_t *a2, int32_t ntime, int32_t nbl,
                            int32_t na, int32_t pb, int32_t qb, int32_t *out, int64_t capacity) {
    using mbx::set_error;
    if (!a1 || !a2 || !out) return set_error(MBX_E_NULL, "build_pair_plan: NULL pointer");
    if (ntime <= 0 || nbl <= 0 || na <= 0 || pb <= 0 || qb <= 0)
        return set_error(MBX_E_DIMS, "build_pair_plan: bad sizes");
    for (int64_t t = 1; t < ntime; ++t)
        for (int64_t b = 0; b < nbl; ++b)
            if (a1[t * nbl + b] != a1[b] || a2[t * nbl + b] != a2[b])
                return set_error(MBX_E_UNSUPPORTED,
                                 "build_pair_plan: antenna pairs change at time %lld", (long long)t);
    const int64_t width = 2 + int64_t(pb) * qb;
    // (P, Q, generation) -> slots
    std::map<std::tuple<int, int, int>, std::vector<int32_t>> blocks;
    for (int32_t b = 0; b < nbl; ++b) {
        const int p = a1[b], q = a2[b];
        if (p < 0 || p >= na || q < 0 || q >= na)
            return set_error(MBX_E_DIMS, "build_pair_plan: baseline %d pairs antennas (%d, %d) "
                             "outside [0, %d)", b, p, q, na);
        const int P = p / pb, Q = q / qb, s = (p % pb) * qb + (q % qb);
        for (int gen = 0;; ++gen) {
            auto &v = blocks[std::make_tuple(P, Q, gen)];
            if (v.empty()) v.assign(size_t(pb) * qb, -1);
            if (v[s] < 0) { v[s] = b; break; }
        }
    }
    if (int64_t(blocks.size()) * width > capacity)
        return set_error(MBX_E_WORKSPACE, "build_pair_plan: needs %lld int32",
                         (long long)(int64_t(blocks.size()) * width));
    int64_t n = 0;
    for (const auto &kv : blocks) {
        int32_t *e = out + n * width;
        e[0] = std::get<0>(kv.first) * pb;
        e[1] = std::get<1>(kv.first) * qb;
        for (size_t i = 0; i < kv.second.size(); ++i) e[2 + i] = kv.second[i];
        ++n;
    }
    return n;
}

int mbx_build_baseline_map(const int32_t *a1, const int32_t *a2, int32_t ntime, int32_t nbl,
                           int32_t na, int32_t *out) {
    using mbx::set_error;
    if (!a1 || !a2 || !out) return set_error(MBX_E_NULL, "build_baseline_map: NULL pointer");
    if (ntime <= 0 || nbl <= 0 || na <= 0) return set_error(MBX_E_DIMS, "build_baseline_map: bad sizes");
    for (int64_t t = 1; t < ntime; ++t)
        for (int64_t b = 0; b < nbl; ++b)
            if (a1[t * nbl + b] != a1[b] || a2[t * nbl + b] != a2[b])
                return set_error(MBX_E_UNSUPPORTED,
                                 "build_baseline_map: antenna pairs change at time %lld", (long long)t);
    for (int64_t i = 0; i < int64_t(na) * na; ++i) out[i] = -1;
    for (int32_t b = 0; b < nbl; ++b) {
        const int p = a1[b], q = a2[b];
        if (p < 0 || p >= na || q < 0 || q >= na)
            return set_error(MBX_E_DIMS, "build_baseline_map: baseline %d pairs (%d, %d) outside "
                             "[0, %d)", b, p, q, na);
        if (out[int64_t(p) * na + q] != -1 || out[int64_t(q) * na + p] != -1)
            return set_error(MBX_E_UNSUPPORTED, "build_baseline_map: antennas (%d, %d) pair "
                             "more than one baseline", p, q);
        if (p == q) {
            out[int64_t(p) * na + p] = b | MBX_BL_AUTO;
        } else {
            out[int64_t(p) * na + q] = b;
            out[int64_t(q) * na + p] = b | MBX_BL_REVERSED;
        }
    }
    return MBX_OK;
}

}  // extern "C"
