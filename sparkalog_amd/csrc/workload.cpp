// Seeded synthetic workloads for the binary recursive operator pipeline.
// Declarations and the contract of every function live in
// include/sparkalog_workload.h. Host-only C++17; built with g++.

#include "sparkalog_workload.h"

#include <algorithm>
#include <parallel/algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kChunkRows = 1ull << 20;

inline uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += kGolden);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t finalize(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Lemire's nearly divisionless bounded integer on [0, range), range >= 1.
inline uint32_t bounded(uint64_t& s, uint32_t range) {
    uint32_t x = static_cast<uint32_t>(splitmix(s) >> 32);
    uint64_t m = static_cast<uint64_t>(x) * range;
    uint32_t l = static_cast<uint32_t>(m);
    if (l < range) {
        const uint32_t t = static_cast<uint32_t>(-range) % range;
        while (l < t) {
            x = static_cast<uint32_t>(splitmix(s) >> 32);
            m = static_cast<uint64_t>(x) * range;
            l = static_cast<uint32_t>(m);
        }
    }
    return static_cast<uint32_t>(m >> 32);
}

// Bounded 64-bit integer on [0, range) by rejection of the biased tail.
inline uint64_t bounded64(uint64_t& s, uint64_t range) {
    if (range <= 0xFFFFFFFFull) return bounded(s, static_cast<uint32_t>(range));
    const uint64_t limit = UINT64_MAX - (UINT64_MAX % range);
    uint64_t x;
    do {
        x = splitmix(s);
    } while (x >= limit);
    return x % range;
}

inline double uniform01(uint64_t& s) {
    return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53;
}

// Knuth's multiplicative Poisson sampler; `limit` = exp(-mean).
inline uint32_t poisson(uint64_t& s, double limit) {
    uint32_t k = 0;
    double p = 1.0;
    do {
        ++k;
        p *= uniform01(s);
    } while (p > limit);
    return k - 1;
}

// Truncated Zipf over the integer sizes [lo, hi] by inverse CDF on the table.
struct SizeTable {
    uint32_t lo;
    std::vector<double> cum;
    SizeTable(double s, uint32_t lo_, uint32_t hi) : lo(lo_) {
        double total = 0.0;
        cum.reserve(hi - lo + 1);
        for (uint32_t k = lo; k <= hi; ++k) {
            const double kd = static_cast<double>(k);
            // k^-1.5 as 1/(k*sqrt(k)) keeps the table free of pow() rounding.
            const double w = (s == 1.5) ? 1.0 / (kd * std::sqrt(kd)) : std::pow(kd, -s);
            total += w;
            cum.push_back(total);
        }
    }
    uint32_t draw(uint64_t& st) const {
        const double u = uniform01(st) * cum.back();
        const auto it = std::upper_bound(cum.begin(), cum.end(), u);
        size_t index = static_cast<size_t>(it - cum.begin());
        if (index >= cum.size()) index = cum.size() - 1;
        return lo + static_cast<uint32_t>(index);
    }
};

// Zipf over ranks 1..n by rejection-inversion (Hoermann & Derflinger 1996).
struct ZipfSampler {
    double s, n, h_x1, h_n, shift;
    ZipfSampler(double exponent, double count) : s(exponent), n(count) {
        h_x1 = H(1.5) - 1.0;
        h_n = H(n + 0.5);
        shift = 2.0 - Hinv(H(2.5) - h(2.0));
    }
    double h(double x) const { return std::exp(-s * std::log(x)); }
    double H(double x) const {
        const double lx = std::log(x);
        return helper2((1.0 - s) * lx) * lx;
    }
    double Hinv(double x) const {
        double t = x * (1.0 - s);
        if (t < -1.0) t = -1.0;
        return std::exp(helper1(t) * x);
    }
    // log1p(x)/x and (exp(x)-1)/x with their limits at 0.
    static double helper1(double x) { return std::fabs(x) > 1e-8 ? std::log1p(x) / x : 1.0 - x * (0.5 - x / 3.0); }
    static double helper2(double x) { return std::fabs(x) > 1e-8 ? std::expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x / 3.0); }
    uint64_t draw(uint64_t& st) const {
        for (;;) {
            const double u = h_n + uniform01(st) * (h_x1 - h_n);
            const double x = Hinv(u);
            double k = std::floor(x + 0.5);
            if (k < 1.0) k = 1.0;
            if (k > n) k = n;
            if (k - x <= shift || u >= H(k + 0.5) - h(k)) return static_cast<uint64_t>(k);
        }
    }
};

inline uint64_t pack(uint32_t a, uint32_t b) { return (static_cast<uint64_t>(a) << 32) | b; }

template <typename F>
void parallel_chunks(uint64_t chunks, F&& body) {
    unsigned threads = std::thread::hardware_concurrency();
    if (threads == 0) threads = 1;
    if (threads > 16) threads = 16;
    if (chunks < threads) threads = static_cast<unsigned>(chunks);
    if (threads <= 1) {
        for (uint64_t c = 0; c < chunks; ++c) body(c);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t) {
        pool.emplace_back([&, t] {
            for (uint64_t c = t; c < chunks; c += threads) body(c);
        });
    }
    for (auto& th : pool) th.join();
}

void draw_pairs(uint64_t seed, uint64_t rows, int key_mode, uint32_t key_space, double zipf_s,
                uint64_t* out) {
    const uint64_t chunks = (rows + kChunkRows - 1) / kChunkRows;
    ZipfSampler zipf(zipf_s > 0 ? zipf_s : 1.1, static_cast<double>(key_space));
    parallel_chunks(chunks, [&](uint64_t c) {
        uint64_t st = finalize(seed ^ (c * kGolden + 0x5A17ull));
        const uint64_t begin = c * kChunkRows;
        const uint64_t end = std::min(rows, begin + kChunkRows);
        for (uint64_t r = begin; r < end; ++r) {
            const uint32_t key = key_mode == 1 ? static_cast<uint32_t>(zipf.draw(st) - 1)
                                               : bounded(st, key_space);
            const uint32_t value = static_cast<uint32_t>(splitmix(st) >> 32);
            out[r] = pack(key, value);
        }
    });
}

// Draw `count` distinct packed pairs (returned sorted).
std::vector<uint64_t> draw_unique_pairs(uint64_t seed, uint64_t count, int key_mode,
                                        uint32_t key_space, double zipf_s) {
    std::vector<uint64_t> pool(count);
    draw_pairs(seed, count, key_mode, key_space, zipf_s, pool.data());
    __gnu_parallel::sort(pool.begin(), pool.end());
    pool.erase(std::unique(pool.begin(), pool.end()), pool.end());
    uint64_t refill = finalize(seed ^ 0xF111ull);
    ZipfSampler zipf(zipf_s > 0 ? zipf_s : 1.1, static_cast<double>(key_space));
    while (pool.size() < count) {
        const uint64_t missing = count - pool.size();
        std::vector<uint64_t> extra(missing);
        for (auto& e : extra) {
            const uint32_t key = key_mode == 1 ? static_cast<uint32_t>(zipf.draw(refill) - 1)
                                               : bounded(refill, key_space);
            e = pack(key, static_cast<uint32_t>(splitmix(refill) >> 32));
        }
        pool.insert(pool.end(), extra.begin(), extra.end());
        __gnu_parallel::sort(pool.begin(), pool.end());
        pool.erase(std::unique(pool.begin(), pool.end()), pool.end());
    }
    return pool;
}

void fisher_yates(uint64_t& st, uint64_t* data, uint64_t n) {
    for (uint64_t i = n; i > 1; --i) {
        const uint64_t j = bounded64(st, i);
        std::swap(data[i - 1], data[j]);
    }
}

// A random permutation for operator inputs of up to 2^28 rows, in parallel:
// position i gets the tag finalize(seed ^ i * kGolden) (distinct for every i,
// as both maps are bijections of u64), and the rows are sorted by tag.
void keyed_shuffle(uint64_t seed, uint64_t* data, uint64_t n) {
    std::vector<std::pair<uint64_t, uint64_t>> tagged(n);
    const uint64_t chunks = (n + kChunkRows - 1) / kChunkRows;
    parallel_chunks(chunks, [&](uint64_t c) {
        const uint64_t end = std::min(n, (c + 1) * kChunkRows);
        for (uint64_t i = c * kChunkRows; i < end; ++i) tagged[i] = {finalize(seed ^ (i * kGolden)), data[i]};
    });
    __gnu_parallel::sort(tagged.begin(), tagged.end());
    parallel_chunks(chunks, [&](uint64_t c) {
        const uint64_t end = std::min(n, (c + 1) * kChunkRows);
        for (uint64_t i = c * kChunkRows; i < end; ++i) data[i] = tagged[i].second;
    });
}

void split_pairs(const uint64_t* packed, uint64_t rows, uint32_t* first, uint32_t* second) {
    for (uint64_t r = 0; r < rows; ++r) {
        first[r] = static_cast<uint32_t>(packed[r] >> 32);
        second[r] = static_cast<uint32_t>(packed[r]);
    }
}

// Random digraph edges of one component [base, base+n) with Poisson(limit) degrees.
void component_edges(uint64_t& st, uint32_t base, uint32_t n, double limit,
                     std::vector<uint64_t>& edges) {
    std::vector<uint32_t> targets;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t degree = poisson(st, limit);
        targets.clear();
        for (uint32_t d = 0; d < degree; ++d) targets.push_back(bounded(st, n));
        std::sort(targets.begin(), targets.end());
        targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
        for (uint32_t t : targets) edges.push_back(pack(base + v, base + t));
    }
}

template <typename T>
T* copy_out(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" uint64_t sparkalog_wl_splitmix64(uint64_t* state) { return splitmix(*state); }

extern "C" void sparkalog_wl_free(void* pointer) { std::free(pointer); }

extern "C" uint64_t sparkalog_wl_mix64(uint64_t key) { return finalize(key + kGolden); }

extern "C" int sparkalog_wl_random_dag(uint64_t seed, uint32_t nodes, uint64_t edges,
                                       uint32_t* src, uint32_t* dst) {
    const uint64_t possible = static_cast<uint64_t>(nodes) * (nodes - 1) / 2;
    if (nodes < 2 || edges > possible) return -1;
    uint64_t st = seed;
    std::vector<uint64_t> chosen;
    chosen.reserve(edges);
    // Draw until `edges` distinct pairs exist; dedupe in batches.
    while (chosen.size() < edges) {
        const uint64_t missing = edges - chosen.size();
        for (uint64_t i = 0; i < missing; ++i) {
            uint32_t a, b;
            do {
                a = bounded(st, nodes);
                b = bounded(st, nodes);
            } while (a == b);
            chosen.push_back(pack(std::min(a, b), std::max(a, b)));
        }
        std::sort(chosen.begin(), chosen.end());
        chosen.erase(std::unique(chosen.begin(), chosen.end()), chosen.end());
    }
    split_pairs(chosen.data(), edges, src, dst);
    return 0;
}

extern "C" uint64_t sparkalog_wl_component_graph(uint64_t seed, uint32_t nodes, double zipf_s,
                                                 uint32_t size_min, uint32_t size_max,
                                                 double mean_out_degree, uint32_t** src,
                                                 uint32_t** dst, uint32_t* components) {
    if (nodes == 0 || size_min == 0 || size_max < size_min) return UINT64_MAX;
    uint64_t st = seed;
    const SizeTable sizes(zipf_s, size_min, size_max);
    std::vector<uint32_t> comp;
    uint64_t used = 0;
    while (used < nodes) {
        uint32_t n = sizes.draw(st);
        if (used + n > nodes) n = static_cast<uint32_t>(nodes - used);
        comp.push_back(n);
        used += n;
    }
    const double limit = std::exp(-mean_out_degree);
    std::vector<uint64_t> edges;
    edges.reserve(static_cast<size_t>(nodes * mean_out_degree * 1.05) + 16);
    uint32_t base = 0;
    for (uint32_t n : comp) {
        component_edges(st, base, n, limit, edges);
        base += n;
    }
    std::vector<uint64_t> perm(nodes);
    for (uint32_t i = 0; i < nodes; ++i) perm[i] = i;
    fisher_yates(st, perm.data(), nodes);
    for (auto& e : edges) {
        e = pack(static_cast<uint32_t>(perm[e >> 32]), static_cast<uint32_t>(perm[e & 0xFFFFFFFFull]));
    }
    __gnu_parallel::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    std::vector<uint32_t> a(edges.size()), b(edges.size());
    split_pairs(edges.data(), edges.size(), a.data(), b.data());
    *src = copy_out(a);
    *dst = copy_out(b);
    if (components) *components = static_cast<uint32_t>(comp.size());
    if (!*src || !*dst) return UINT64_MAX;
    return edges.size();
}

extern "C" uint64_t sparkalog_wl_batch_graphs(uint64_t seed, uint32_t instances, uint32_t n_min,
                                              uint32_t n_max, double lambda_min,
                                              double lambda_max, uint32_t** src, uint32_t** dst,
                                              uint64_t** offsets, uint32_t** node_counts) {
    if (instances == 0 || n_min == 0 || n_max < n_min) return UINT64_MAX;
    uint64_t st = seed;
    std::vector<uint32_t> counts(instances);
    std::vector<double> lambdas(instances);
    const double log_lo = std::log(static_cast<double>(n_min));
    const double log_hi = std::log(static_cast<double>(n_max) + 1.0);
    for (uint32_t i = 0; i < instances; ++i) {
        double n = std::floor(std::exp(log_lo + uniform01(st) * (log_hi - log_lo)));
        if (n < n_min) n = n_min;
        if (n > n_max) n = n_max;
        counts[i] = static_cast<uint32_t>(n);
        lambdas[i] = lambda_min + uniform01(st) * (lambda_max - lambda_min);
    }
    // Instances draw from independent sub-streams so they can be built in parallel.
    std::vector<std::vector<uint64_t>> per(instances);
    parallel_chunks(instances, [&](uint64_t i) {
        uint64_t local = finalize(seed ^ ((i + 1) * 0xD1B54A32D192ED03ull));
        component_edges(local, 0, counts[i], std::exp(-lambdas[i]), per[i]);
        std::sort(per[i].begin(), per[i].end());
    });
    std::vector<uint64_t> offs(instances + 1, 0);
    for (uint32_t i = 0; i < instances; ++i) offs[i + 1] = offs[i] + per[i].size();
    const uint64_t total = offs[instances];
    uint32_t* a = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(1, total) * 4));
    uint32_t* b = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(1, total) * 4));
    if (!a || !b) {
        std::free(a);
        std::free(b);
        return UINT64_MAX;
    }
    for (uint32_t i = 0; i < instances; ++i) split_pairs(per[i].data(), per[i].size(), a + offs[i], b + offs[i]);
    *src = a;
    *dst = b;
    *offsets = copy_out(offs);
    *node_counts = copy_out(counts);
    return total;
}

extern "C" void sparkalog_wl_pairs(uint64_t seed, uint64_t rows, int key_mode, uint32_t key_space,
                                   double zipf_s, uint32_t* first, uint32_t* second) {
    std::vector<uint64_t> packed(rows);
    draw_pairs(seed, rows, key_mode, key_space, zipf_s, packed.data());
    split_pairs(packed.data(), rows, first, second);
}

extern "C" void sparkalog_wl_distinct_input(uint64_t seed, uint64_t rows, double unique_fraction,
                                            int key_mode, uint32_t key_space, double zipf_s,
                                            uint32_t* first, uint32_t* second) {
    if (rows == 0) return;
    uint64_t unique = static_cast<uint64_t>(std::llround(unique_fraction * static_cast<double>(rows)));
    if (unique < 1) unique = 1;
    if (unique > rows) unique = rows;
    std::vector<uint64_t> pool = draw_unique_pairs(seed, unique, key_mode, key_space, zipf_s);
    pool.resize(rows);
    uint64_t st = finalize(seed ^ 0xD0B1Eull);
    for (uint64_t r = unique; r < rows; ++r) pool[r] = pool[bounded64(st, unique)];
    keyed_shuffle(finalize(st), pool.data(), rows);
    split_pairs(pool.data(), rows, first, second);
}

extern "C" void sparkalog_wl_overlapping_sets(uint64_t seed, uint64_t rows, double overlap, int key_mode,
                                              uint32_t key_space, double zipf_s,
                                              uint32_t* left_first, uint32_t* left_second,
                                              uint32_t* right_first, uint32_t* right_second) {
    if (rows == 0) return;
    uint64_t shared = static_cast<uint64_t>(std::llround(overlap * static_cast<double>(rows)));
    if (shared > rows) shared = rows;
    const uint64_t total = 2 * rows - shared;
    const std::vector<uint64_t> pool = draw_unique_pairs(seed, total, key_mode, key_space, zipf_s);
    // One pass over the sorted pool assigns every pair to both sides, the
    // left only or the right only, with probabilities proportional to the
    // quotas still open (sequential selection sampling, Knuth's Algorithm S):
    // exact counts, a uniformly random split, and both sides come out sorted.
    uint64_t st = finalize(seed ^ 0x0E71Aull);
    uint64_t open_shared = shared, open_left = rows - shared, open_right = rows - shared;
    uint64_t nl = 0, nr = 0;
    for (uint64_t i = 0; i < total; ++i) {
        const uint64_t u = bounded64(st, open_shared + open_left + open_right);
        const uint64_t k = pool[i];
        if (u < open_shared) {
            --open_shared;
            left_first[nl] = static_cast<uint32_t>(k >> 32);
            left_second[nl++] = static_cast<uint32_t>(k);
            right_first[nr] = static_cast<uint32_t>(k >> 32);
            right_second[nr++] = static_cast<uint32_t>(k);
        } else if (u < open_shared + open_left) {
            --open_left;
            left_first[nl] = static_cast<uint32_t>(k >> 32);
            left_second[nl++] = static_cast<uint32_t>(k);
        } else {
            --open_right;
            right_first[nr] = static_cast<uint32_t>(k >> 32);
            right_second[nr++] = static_cast<uint32_t>(k);
        }
    }
}

extern "C" uint64_t sparkalog_wl_pair_digest(const uint32_t* first, const uint32_t* second,
                                             uint64_t rows) {
    uint64_t sum = 0;
    for (uint64_t r = 0; r < rows; ++r) sum += finalize(pack(first[r], second[r]) + kGolden);
    return sum;
}

extern "C" uint64_t sparkalog_wl_packed_digest(const uint64_t* packed, uint64_t rows) {
    uint64_t sum = 0;
    for (uint64_t r = 0; r < rows; ++r) sum += finalize(packed[r] + kGolden);
    return sum;
}

namespace {

char* put_u32(char* p, uint32_t v) {
    char buf[10];
    int n = 0;
    do {
        buf[n++] = static_cast<char>('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) *p++ = buf[--n];
    return p;
}

}  // namespace

extern "C" uint64_t sparkalog_wl_edge_text(uint64_t seed, uint64_t rows, uint32_t key_space, uint32_t style,
                                           char** text, uint32_t* first, uint32_t* second) {
    // at most 80 bytes per row (two 10-digit values with '+' and up to 3
    // leading zeros, separators, padding, CRLF) plus a comment or blank line
    const uint64_t cap = rows * 80 + 64;
    char* out = static_cast<char*>(std::malloc(cap ? cap : 1));
    if (!out) return UINT64_MAX;
    uint64_t s = seed;
    char* p = out;
    const char* seps[] = {" ", "\t", "  ", " \t "};
    for (uint64_t r = 0; r < rows; ++r) {
        const uint64_t bits = splitmix(s);
        if ((style & 2) && (bits & 63) == 0) {
            const char* extra[] = {"# comment 1 2 3", "% note", "", "   \t"};
            const char* e = extra[(bits >> 6) & 3];
            while (*e) *p++ = *e++;
            *p++ = '\n';
        }
        const uint32_t a = key_space ? bounded(s, key_space) : static_cast<uint32_t>(splitmix(s) >> 32);
        const uint32_t b = key_space ? bounded(s, key_space) : static_cast<uint32_t>(splitmix(s) >> 32);
        first[r] = a;
        second[r] = b;
        if ((style & 1) && ((bits >> 8) & 7) == 0) *p++ = ' ';
        if ((style & 8) && ((bits >> 11) & 7) == 0) *p++ = '+';
        if ((style & 8) && ((bits >> 14) & 7) == 0) *p++ = '0';
        p = put_u32(p, a);
        const char* sep = (style & 1) ? seps[(bits >> 17) & 3] : " ";
        while (*sep) *p++ = *sep++;
        if ((style & 8) && ((bits >> 19) & 7) == 0) *p++ = '+';
        p = put_u32(p, b);
        if ((style & 1) && ((bits >> 22) & 7) == 0) *p++ = '\t';
        if ((style & 4) && ((bits >> 25) & 1)) *p++ = '\r';
        if (r + 1 < rows || !(style & 16)) *p++ = '\n';
    }
    *text = out;
    return static_cast<uint64_t>(p - out);
}
