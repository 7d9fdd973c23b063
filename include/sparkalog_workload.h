/*
 * Seeded synthetic workloads for the binary recursive operator pipeline.
 *
 * Every generator is deterministic: the same seed and sizes produce the same
 * arrays on every machine that runs this library's build. Random numbers come
 * from SplitMix64 (Steele, Lea & Flood 2014) with Lemire's bounded integers
 * (Lemire 2019). Generators that need a large output allocate it with malloc
 * and hand ownership to the caller, who releases it with sparkalog_wl_free.
 *
 * The reference has no generator for these shapes: its tc-fixpoint driver
 * builds a chain (crates/sparkalog-cli/src/bin/tc-fixpoint.rs:120-125) and its
 * crossover harnesses read SNAP com-dblp from disk. These workloads restate
 * the five benchmark configurations of the binary pipeline as seeded draws.
 */
#ifndef SPARKALOG_WORKLOAD_H
#define SPARKALOG_WORKLOAD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One SplitMix64 step; exposed so that callers can derive sub-seeds. */
uint64_t sparkalog_wl_splitmix64(uint64_t* state);

/* Release an array returned by one of the generators below. */
void sparkalog_wl_free(void* pointer);

/*
 * Random DAG: draw ordered pairs (a, b), a != b, uniform on [0, nodes); keep
 * (min, max); repeat until `edges` distinct pairs exist. Writes the pairs
 * sorted lexicographically into caller buffers of `edges` entries.
 * Returns 0, or -1 when `edges` exceeds nodes*(nodes-1)/2.
 */
int sparkalog_wl_random_dag(uint64_t seed, uint32_t nodes, uint64_t edges,
                            uint32_t* src, uint32_t* dst);

/*
 * Component graph: `nodes` ids split into components whose sizes follow a
 * Zipf(zipf_s) law truncated to [size_min, size_max] (inverse CDF over the
 * size table; the last component is clipped to fit). Each node draws a
 * Poisson(mean_out_degree) out-degree (Knuth's method) with targets uniform
 * inside its component; duplicate edges are dropped, self-loops kept. Node ids
 * are then permuted by Fisher-Yates. Output: sorted unique (src, dst) pairs in
 * malloc'd arrays; *components receives the component count.
 * Returns the number of edges, or UINT64_MAX on allocation failure.
 */
uint64_t sparkalog_wl_component_graph(uint64_t seed, uint32_t nodes, double zipf_s,
                                      uint32_t size_min, uint32_t size_max,
                                      double mean_out_degree, uint32_t** src,
                                      uint32_t** dst, uint32_t* components);

/*
 * Batch of independent digraphs: `instances` graphs, node count log-uniform on
 * [n_min, n_max], Poisson out-degree with mean uniform on [lambda_min,
 * lambda_max] drawn per instance, uniform targets, local ids 0..n-1, duplicates
 * dropped, self-loops kept. Edges of instance i occupy rows
 * [(*offsets)[i], (*offsets)[i+1]) sorted by (src, dst); (*node_counts)[i] is
 * its node count. Returns the total edge count, or UINT64_MAX on failure.
 */
uint64_t sparkalog_wl_batch_graphs(uint64_t seed, uint32_t instances, uint32_t n_min,
                                   uint32_t n_max, double lambda_min, double lambda_max,
                                   uint32_t** src, uint32_t** dst, uint64_t** offsets,
                                   uint32_t** node_counts);

/*
 * Operator inputs. `rows` pairs written to caller buffers.
 *   key_mode 0: first column uniform on [0, key_space)
 *   key_mode 1: first column Zipf(zipf_s) over key_space ranks (rank r -> key r-1)
 * The second column is uniform on [0, 2^32). Generated in fixed chunks of
 * 2^20 rows with per-chunk seeds, so the result does not depend on threads.
 */
void sparkalog_wl_pairs(uint64_t seed, uint64_t rows, int key_mode, uint32_t key_space,
                        double zipf_s, uint32_t* first, uint32_t* second);

/*
 * Distinct input: round(unique_fraction*rows) distinct pairs drawn as by
 * sparkalog_wl_pairs (redrawn on collision), the remaining rows resampled
 * uniformly from them, then the whole array shuffled by Fisher-Yates.
 */
void sparkalog_wl_distinct_input(uint64_t seed, uint64_t rows, double unique_fraction,
                                 int key_mode, uint32_t key_space, double zipf_s,
                                 uint32_t* first, uint32_t* second);

/*
 * Two sorted unique relations of `rows` pairs each that share
 * round(overlap*rows) pairs: a pool of 2*rows - shared unique pairs (first
 * column uniform on [0, key_space) for key_mode 0, Zipf(zipf_s) ranks for
 * key_mode 1; second column uniform) is drawn, shuffled, split, and each side
 * sorted.
 */
void sparkalog_wl_overlapping_sets(uint64_t seed, uint64_t rows, double overlap, int key_mode,
                                   uint32_t key_space, double zipf_s,
                                   uint32_t* left_first, uint32_t* left_second,
                                   uint32_t* right_first, uint32_t* right_second);

/* Order-independent 64-bit digest of a pair set: sum of mix64((a<<32)|b). */
uint64_t sparkalog_wl_pair_digest(const uint32_t* first, const uint32_t* second,
                                  uint64_t rows);

/* The same digest over packed keys (a<<32)|b. */
uint64_t sparkalog_wl_packed_digest(const uint64_t* packed, uint64_t rows);

/* The mixing function of the digest (the SplitMix64 finaliser). */
uint64_t sparkalog_wl_mix64(uint64_t key);

/*
 * Text of `rows` binary u32 facts in the reference loader's format
 * (crates/sparkalog-storage/src/io.rs:72-141), with the values written to
 * first/second (caller buffers of `rows` entries) as the ground truth of a
 * parse. Values are uniform on [0, key_space) (0 = the whole u32 range).
 * style bits: 1 = mixed whitespace (tabs, runs, leading/trailing), 2 = a
 * comment or blank line before about one row in 64, 4 = CRLF on about half
 * the lines, 8 = '+' signs and leading zeros, 16 = no newline after the
 * last row. *text is malloc'd (release with sparkalog_wl_free); returns its
 * length, or UINT64_MAX on allocation failure.
 */
uint64_t sparkalog_wl_edge_text(uint64_t seed, uint64_t rows, uint32_t key_space, uint32_t style,
                                char** text, uint32_t* first, uint32_t* second);

#ifdef __cplusplus
}
#endif

#endif
