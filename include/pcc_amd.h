/*
 * pcc_amd.h -- C ABI of the MI355X-native ParallelCorrelationClusterer /
 * ParallelModularityClusterer hot path.
 *
 * The library keeps the graph CSR and the clustering state resident in HBM for
 * the whole Cluster call and runs the Louvain best-move rounds, the cluster
 * contraction and the multilevel refinement as hand-written gfx950 kernels.
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).
 *
 * Conventions:
 *   - Return values are absl::StatusCode numbers (PCC_OK ... PCC_INTERNAL);
 *     pcc_last_error() gives the thread-local message of the last failure.
 *   - The caller owns every host array.  A pcc_graph owns its device buffers
 *     and one HIP stream on the device that was current when it was created.
 *   - Handles are independent; a single handle is not reentrant.
 *   - Vertex ids and cluster ids are uint32 (gbbs::uintE); n < 2^32 - 1.
 */
#ifndef PCC_AMD_H_
#define PCC_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* absl::StatusCode values used by the reference (clustering/status_macros.h). */
#define PCC_OK 0
#define PCC_INVALID_ARGUMENT 3
#define PCC_NOT_FOUND 5
#define PCC_RESOURCE_EXHAUSTED 8
#define PCC_UNIMPLEMENTED 12
#define PCC_INTERNAL 13

/* Sentinel for "no move" in per-vertex move arrays. */
#define PCC_NO_MOVE 0xFFFFFFFFu

/* CorrelationClustererConfig enums, clustering/config.proto:115-160. */
#define PCC_DEFAULT_CLUSTER_MOVES 0
#define PCC_LOUVAIN 1
#define PCC_NBHR_CLUSTER_MOVE 0
#define PCC_ALL_MOVE 1
#define PCC_NBHR_MOVE 2

/* Objective selector: ParallelCorrelationClusterer or
 * ParallelModularityClusterer (clustering/cluster-in-memory_main.cc:270-275). */
#define PCC_OBJECTIVE_CORRELATION 0
#define PCC_OBJECTIVE_MODULARITY 1

/*
 * CorrelationClustererConfig (clustering/config.proto:94-183) with its proto2
 * defaults.  Fields the path does not read are kept so that every text-format
 * config the reference accepts parses here too.
 */
typedef struct pcc_config {
  double resolution;                     /* 1 */
  double edge_weight_offset;             /* 2 */
  uint32_t random_seed;                  /* 3 (not read by the parallel path) */
  int64_t num_iterations;                /* 4 */
  int64_t louvain_num_iterations;        /* 8.1 LouvainConfig.num_iterations */
  int64_t louvain_num_inner_iterations;  /* 8.2 LouvainConfig.num_inner_iterations */
  int32_t clustering_moves_method;       /* 9, default PCC_LOUVAIN */
  int32_t subclustering_method;          /* 10: 0 NONE, 1 TRIANGLE, 2 CONNECTIVITY, 3 SLOW_CONNECTIVITY (batches: graph by graph) */
  int32_t preclustering_method;          /* 11, only NONE_PRECLUSTERING (0) */
  int32_t move_method;                   /* 16, default PCC_NBHR_CLUSTER_MOVE */
  uint8_t refine;                        /* 13 */
  uint8_t async;                         /* 14 */
  uint8_t connect_stats;                 /* 15 (never read by the reference) */
  uint8_t all_iter;                      /* 17 (sequential clusterers only) */
  uint8_t permute;                       /* 18 */
  uint8_t _pad[3];
  uint64_t permute_seed;                 /* processing-order seed (async + permute) */
} pcc_config;

/* Proto2 defaults of CorrelationClustererConfig. */
void pcc_config_default(pcc_config* cfg);

/* Parses a text-format ClustererConfig holding a correlation_clusterer_config
 * (e.g. "correlation_clusterer_config { resolution: 0.01 move_method: ALL_MOVE }")
 * or the bare CorrelationClustererConfig fields.  Replaces
 * google::protobuf::TextFormat::ParseFromString at
 * clustering/cluster-in-memory_main.cc:257-265: PCC_INVALID_ARGUMENT on
 * malformed text or unknown fields. */
int pcc_config_parse_text(const char* text, pcc_config* out);

/* ---------------------------------------------------------------- graph -- */

typedef struct pcc_graph pcc_graph;

/* New empty graph under construction; replaces GbbsGraph's constructor
 * (clustering/gbbs-graph.h:39-56). */
int pcc_graph_new(pcc_graph** out);

/* InMemoryClusterer::Graph::Import (clustering/in-memory-clusterer.h:60-65,
 * clustering/gbbs-graph.cc:32-49).  Safe to call concurrently from several
 * threads on one graph.  node_weight is accepted and ignored, as GbbsGraph
 * ignores it (clustering/gbbs-graph.h:34).  weights may be NULL (1.0). */
int pcc_graph_import(pcc_graph* g, uint32_t id, double node_weight,
                     const uint32_t* nbrs, const double* weights, uint64_t degree);

/* GbbsGraph::FinishImport (clustering/gbbs-graph.cc:51-76): builds the CSR in
 * import order, rounds weights to float32, and uploads it to HBM. */
int pcc_graph_finish_import(pcc_graph* g);

/* Direct construction from a host CSR, the path the reference's CLI takes
 * (clustering/cluster-in-memory_main.cc:231-250 CopyGraph + :306).  offsets
 * has n+1 entries; weights may be NULL (unit weights).  The graph must be
 * symmetric (clustering/clusterers/parallel-correlation-clusterer.h:43). */
int pcc_graph_from_csr(int64_t n, int64_t m, const int64_t* offsets,
                       const uint32_t* nbrs, const float* weights, pcc_graph** out);

int64_t pcc_graph_num_nodes(const pcc_graph* g);
int64_t pcc_graph_num_edges(const pcc_graph* g);

/* Copies the device CSR back to host arrays (n+1, m, m entries). */
int pcc_graph_copy_csr(const pcc_graph* g, int64_t* offsets, uint32_t* nbrs, float* weights);

void pcc_graph_destroy(pcc_graph* g);

/* ---------------------------------------------------------------- cluster -- */

typedef struct pcc_stats {
  int64_t levels;              /* levels run (final iter + 1) */
  int64_t rounds;              /* best-move rounds, including refinement */
  int64_t edges_scanned;       /* sum over rounds of sum_{v in frontier} deg(v) */
  int64_t frontier_vertices;   /* sum over rounds of |frontier| */
  int64_t num_clusters;        /* clusters in the output */
  int64_t fallback_vertices;   /* vertices whose sums took the ordered path */
  double objective;            /* ComputeObjective on the input graph */
  double modularity;           /* ComputeModularity (modularity runs) */
  double cluster_seconds;      /* span of "Cluster Time:", device-side */
  double bestmove_seconds;     /* sum of best-move kernel time (timing on) */
  double bestmove_bytes;       /* algorithmic bytes 12*TE + 40*|F| summed */
  double contract_seconds;     /* contraction time (timing on) */
  double apply_seconds;        /* apply + frontier time (timing on) */
  int64_t bestmove_launches;   /* best-move kernel launches (timing on) */
  double ordered_seconds;      /* ordered fallback kernel time (timing on) */
  /* Per degree bin (0..4: deg <= 4, 8, 16, 32, 64; 5: one-wave LDS hash;
   * 6: one-workgroup LDS hash; 7: global-memory hash). */
  double bin_seconds[8];       /* kernel time (timing on) */
  int64_t bin_edges[8];        /* edges scanned */
  int64_t bin_vertices[8];     /* vertices processed */
  int64_t bin_launches[8];     /* launches (timing on) */
  /* Whole-cluster phase (clustering_moves_method DEFAULT_CLUSTER_MOVES). */
  double cluster_move_seconds; /* kernel time (timing on) */
  int64_t cluster_moves;       /* clusters moved */
} pcc_stats;

/* ParallelCorrelationClusterer::Cluster
 * (clustering/clusterers/parallel-correlation-clusterer.cc:641-653) or, with
 * objective == PCC_OBJECTIVE_MODULARITY, ParallelModularityClusterer::Cluster
 * (clustering/clusterers/parallel-modularity-clusterer.cc:98-111).
 * out_cluster_ids[n] receives the reference's final cluster_ids vector
 * (before OutputIndicesById, parallel-correlation-clusterer.cc:606-616).
 * stats may be NULL. */
int pcc_cluster(pcc_graph* g, const pcc_config* cfg, int objective,
                uint32_t* out_cluster_ids, pcc_stats* stats);

/* RefineClusters(config, clustering, node_weights)
 * (parallel-correlation-clusterer.cc:399-412, 631-639;
 * parallel-modularity-clusterer.cc:68-96).  inout_cluster_ids[n] holds the
 * initial clustering as per-vertex cluster indices in [0, n) (cluster i of the
 * reference's Clustering vector, its members taken in ascending order); it is
 * overwritten with the result.  node_weights may be NULL (1.0, or degrees for
 * modularity).  Any finite weights are accepted: when they are not all
 * integers (or their magnitudes sum to 2^53 or more) the cluster weights are
 * updated in the reference's order instead of by atomics, so sync results stay
 * exact; DEFAULT_CLUSTER_MOVES then returns PCC_INVALID_ARGUMENT. */
int pcc_refine_clusters(pcc_graph* g, const pcc_config* cfg, int objective,
                        const double* node_weights, uint32_t* inout_cluster_ids,
                        pcc_stats* stats);

/* ClusteringHelper::ComputeObjective
 * (parallel-correlation-clusterer-internal.cc:118-149) on device. */
int pcc_compute_objective(pcc_graph* g, const pcc_config* cfg,
                          const uint32_t* cluster_ids, const double* node_weights,
                          double* out_objective);

/* No counterpart in the reference: pcc_compute_objective for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same per-vertex terms and
 * the same fixed-order sums give the same value bit for bit (when every
 * cluster weight is an exact sum of node weights, e.g. integer weights;
 * fractional weights are summed by atomic adds in an unspecified order), but
 * no device memory is allocated
 * and nothing synchronizes: cluster_ids_dev (n entries) and node_weights_dev
 * (n entries, NULL: unit weights) are device arrays, *out_objective_dev a
 * device double written on `stream` (a hipStream_t; NULL: the handle's
 * stream).  The work memory is one fixed array per device (about 50 MiB,
 * kept until the process ends), so the graph may have at most 2^21 vertices
 * and 2^30 directed edges (PCC_INVALID_ARGUMENT beyond).  Cluster ids
 * outside [0, n) make the objective NaN.  Calls that overlap on one device
 * share that array and race: serialize them (one stream, or events). */
int pcc_compute_objective_noalloc(pcc_graph* g, const pcc_config* cfg,
                                  const uint32_t* cluster_ids_dev,
                                  const double* node_weights_dev,
                                  double* out_objective_dev, void* stream);

/* ComputeModularity (parallel-modularity-clusterer.cc:36-66) on device. */
int pcc_compute_modularity(pcc_graph* g, double resolution,
                           const uint32_t* cluster_ids, double* out_modularity);

/* ---------------------------------------------------------------- batch -- */
/* Many independent graphs clustered together: every phase of every graph runs
 * in the same kernel launches (one launch per phase and bin for the whole
 * batch instead of one per graph), while each graph keeps the control flow,
 * id numbering and results of its own pcc_cluster run.  No counterpart in the
 * reference, whose Cluster takes one graph; the per-graph semantics are those
 * of ParallelCorrelationClusterer::Cluster / ParallelModularityClusterer::Cluster.
 *
 * Graph g has vertex_counts[g] >= 1 vertices; the CSR rows of all graphs are
 * concatenated (offsets has sum(vertex_counts) + 1 entries, offsets[0] = 0)
 * and each graph's neighbour ids are local to that graph. */
int pcc_batch_from_csr(int64_t num_graphs, const int64_t* vertex_counts,
                       const int64_t* offsets, const uint32_t* nbrs,
                       const float* weights, pcc_graph** out);
int64_t pcc_batch_num_graphs(const pcc_graph* g);

/* Clusters every graph of a batch.  resolutions[g] (NULL: cfg->resolution)
 * is graph g's resolution (its gamma for modularity); the other config fields
 * apply to all graphs.  out_cluster_ids receives, for every vertex, its final
 * cluster id local to its graph (what pcc_cluster returns for that graph
 * alone).  out_objectives / out_modularities (may be NULL) receive one value
 * per graph. */
int pcc_batch_cluster(pcc_graph* batch, const pcc_config* cfg, int objective,
                      const double* resolutions, uint32_t* out_cluster_ids,
                      double* out_objectives, double* out_modularities, pcc_stats* stats);

/* -------------------------------------------------------------- serving -- */
/* A request loop over the batch path (no counterpart in the reference, which
 * clusters one graph per call; for callers that serve clustering requests).
 * Submitted graphs queue on the host; worker threads (one HIP stream each)
 * repeatedly form a batch from the queue, admitting requests while the batch's
 * estimated HBM footprint fits the budget left by the batches already on the
 * device, so requests that arrive while others run are picked up as soon as a
 * worker and memory free up, and batches on different workers overlap on the
 * GPU.  Each result equals that request's pcc_cluster run. */
typedef struct pcc_server pcc_server;

typedef struct pcc_server_options {
  int32_t workers;              /* batches in flight at once (default 2) */
  int32_t _pad;
  int64_t max_batch_graphs;     /* default 512 */
  int64_t max_batch_edges;      /* default 2^31 directed edges */
  double hbm_budget_bytes;      /* admission budget; 0: 90% of free HBM at creation */
  double batch_window_seconds;  /* wait this long for more requests before a batch (default 0) */
} pcc_server_options;

typedef struct pcc_server_stats {
  int64_t completed;
  int64_t failed;
  int64_t batches;
  int64_t max_graphs_in_flight;
  double p50_latency_seconds;
  double p99_latency_seconds;
  double mean_batch_graphs;
  double graphs_per_second;     /* completed / (last completion - first submission) */
} pcc_server_stats;

void pcc_server_options_default(pcc_server_options* opt);
/* A config with subclustering_method != 0 runs every request as a batch of
 * its own graph (max_batch_graphs is taken as 1). */
int pcc_server_create(const pcc_config* cfg, int objective, const pcc_server_options* opt,
                      pcc_server** out);
/* Copies the graph (local ids, n >= 1) and queues it; returns its request id. */
int pcc_server_submit(pcc_server* s, int64_t n, int64_t m, const int64_t* offsets,
                      const uint32_t* nbrs, const float* weights, double resolution,
                      uint64_t* out_request_id);
/* Blocks until the request is done; fills its n cluster ids (local), objective
 * and submit-to-completion latency.  Returns the request's status.  A request
 * is claimed by its first wait: another wait on the same id (concurrent or
 * later) returns PCC_NOT_FOUND. */
int pcc_server_wait(pcc_server* s, uint64_t request_id, uint32_t* out_cluster_ids,
                    double* out_objective, double* out_latency_seconds);
int pcc_server_get_stats(pcc_server* s, pcc_server_stats* out);
void pcc_server_destroy(pcc_server* s);

/* ------------------------------------------------------------ one phase -- */
/* Phase-level entry points used by parity tests and by callers that drive the
 * Louvain loop themselves.  State arrays are the ClusteringHelper SoA
 * (parallel-correlation-clusterer-internal.h:158-164), n entries each. */

/* The sync vertex map of BestMovesForVertexSubset
 * (parallel-correlation-clusterer.cc:132-160): EfficientBestMove
 * (parallel-correlation-clusterer-internal.cc:448-583), the singleton rule and
 * the delta > 0 filter, for every vertex with frontier[v] != 0 (frontier NULL:
 * all).  out_moves[v] = target or PCC_NO_MOVE, out_moved[v] in {0,1};
 * out_target / out_delta (may be NULL) hold EfficientBestMove's own return. */
int pcc_best_moves(pcc_graph* g, const pcc_config* cfg, const uint32_t* cluster_ids,
                   const uint32_t* cluster_sizes, const double* node_weights,
                   const double* cluster_weights, const uint8_t* frontier,
                   uint32_t* out_moves, uint8_t* out_moved, uint32_t* out_target,
                   double* out_delta);

/* One BestMovesForVertexSubset round (parallel-correlation-clusterer.cc:115-338)
 * in place: state updated as MoveNodesToCluster / AsyncMove leave it, frontier
 * (n bytes) replaced by the next frontier, whose size is returned in
 * out_frontier_size.  Config checks and node-weight handling as in
 * pcc_refine_clusters. */
int pcc_round(pcc_graph* g, const pcc_config* cfg, uint32_t* cluster_ids,
              uint32_t* cluster_sizes, const double* node_weights,
              double* cluster_weights, uint8_t* frontier, int64_t* out_frontier_size);

/* CompressGraph (parallel-correlation-clusterer-internal.cc:709-782):
 * creates *out_graph = the contracted graph and writes its aggregated node
 * weights to out_node_weights (capacity n; n' = pcc_graph_num_nodes(*out_graph)
 * entries are written). */
int pcc_compress(pcc_graph* g, const uint32_t* cluster_ids, const double* node_weights,
                 pcc_graph** out_graph, double* out_node_weights);

/* --------------------------------------------------------------- tuning -- */

/* Runtime options of a graph handle.  Keys: */
#define PCC_OPT_TIMING 1         /* 1: time kernels with HIP events into pcc_stats */
#define PCC_OPT_FORCE_ORDERED 2  /* 1: route every hashed vertex to the ordered path */
#define PCC_OPT_WAVE_MAX_DEG 3   /* upper degree of the one-wave LDS-hash bin */
#define PCC_OPT_BLOCK_MAX_DEG 4  /* upper degree of the one-workgroup LDS-hash bin */
#define PCC_OPT_ASYNC_SUBROUNDS 5 /* async: launches per round, each on 1/k of the frontier (0: automatic) */
#define PCC_OPT_ASYNC_SINGLETON_RULE 6 /* async: 1 applies the sync singleton rule (default 0) */
#define PCC_OPT_HUB_PART_EDGES 7 /* edges per hash partition of the largest-degree bin */
#define PCC_OPT_PULL_FRACTION 8  /* next-frontier expansion pulls when the seeds' (degree + 1)
                                    sum exceeds value/1024 of (m + n); symmetric graphs only;
                                    1024 disables (default 512) */
#define PCC_OPT_HASH_WAVES 10      /* workgroups per hash-bin launch as a multiple of those
                                       resident at once (default 1; 0: one per vertex up to
                                       8192) */
#define PCC_OPT_CLUSTER_SUBROUNDS 9 /* async whole-cluster moves: sub-rounds per round
                                       (default 16, range [1, 1024]) */
#define PCC_OPT_DEVICE_ROUNDS 11   /* 1: sync LOUVAIN rounds with their counts on device,
                                      one round captured as a HIP graph and replayed when
                                      PCC_OPT_TIMING is off; 2: the same rounds launched one
                                      by one; 0 (default, measured fastest): the host reads
                                      every round's frontier counts back.  All three give
                                      the same results. */
#define PCC_OPT_CONTRACT_METHOD 12 /* contraction: 0 (default) cluster-grouped sorts on chip;
                                       1 one LSD radix sort of every kept edge.  Both give
                                       the same contracted graph. */
#define PCC_OPT_CONTRACT_CHUNK 13  /* cluster-grouped contraction: most terms one radix sort of
                                      its large clusters takes (default and maximum 2^31);
                                      smaller values bound the sort's temporary memory */
int pcc_set_option(pcc_graph* g, int key, int64_t value);

/* Autotuner of the degree-bin thresholds (PCC_OPT_WAVE_MAX_DEG,
 * PCC_OPT_BLOCK_MAX_DEG) and the hub partition size (PCC_OPT_HUB_PART_EDGES)
 * for one workload: runs whole Cluster calls on the handle (a single graph or
 * a batch; resolutions per graph as in pcc_batch_cluster, may be NULL) for
 * candidate settings by coordinate descent from the current ones, timing the
 * device span of each call, and leaves the fastest setting on the handle.
 * These settings only choose which kernel handles a vertex: the clustering
 * itself does not depend on them.  max_trials <= 0 means the default budget
 * (one sweep over the three settings).  No counterpart in the reference (its
 * CPU path has no kernel shapes to choose); it serves the batched and serving
 * forms. */
typedef struct pcc_tune_result {
  int64_t wave_max_deg;      /* chosen settings */
  int64_t block_max_deg;
  int64_t hub_part_edges;
  int64_t trials;            /* Cluster calls made */
  double initial_seconds;    /* device span with the settings found on entry */
  double best_seconds;       /* device span with the chosen settings */
} pcc_tune_result;
int pcc_autotune(pcc_graph* g, const pcc_config* cfg, int objective,
                 const double* resolutions, int max_trials, pcc_tune_result* out);

/* ------------------------------------------------------------ host data -- */

/* Host CSR returned by the generators and the GBBS adjacency reader; release
 * with pcc_host_csr_free. */
typedef struct pcc_host_csr {
  int64_t n;
  int64_t m;
  int64_t* offsets;   /* n + 1 */
  uint32_t* nbrs;     /* m */
  float* weights;     /* m */
} pcc_host_csr;

void pcc_host_csr_free(pcc_host_csr* csr);

/* Planted-partition SBM: num_blocks equal blocks (the last takes the
 * remainder), undirected, no self-loops, pair probabilities p_in / p_out,
 * weights uniform in [w_lo, w_hi] (w_lo == w_hi: constant), float32-rounded.
 * Deterministic in seed. */
int pcc_gen_sbm(int64_t n, int64_t num_blocks, double p_in, double p_out,
                double w_lo, double w_hi, uint64_t seed, pcc_host_csr* out);

/* SBM with explicit block sizes (sum = n). */
int pcc_gen_sbm_blocks(int64_t num_blocks, const int64_t* block_sizes, double p_in,
                       double p_out, double w_lo, double w_hi, uint64_t seed,
                       pcc_host_csr* out);

/* RMAT: 2^scale vertices, edge_factor * 2^scale draws with quadrant
 * probabilities a, b, c (d = 1 - a - b - c), symmetrized, deduplicated,
 * self-loops removed; weights uniform in [w_lo, w_hi] per undirected pair. */
int pcc_gen_rmat(int scale, double edge_factor, double a, double b, double c,
                 double w_lo, double w_hi, uint64_t seed, pcc_host_csr* out);

/* GBBS AdjacencyGraph / WeightedAdjacencyGraph text format, the reference's
 * input (read_unweighted_symmetric_graph / read_weighted_symmetric_graph at
 * clustering/cluster-in-memory_main.cc:292-305). */
int pcc_read_adjacency_graph(const char* path, int float_weighted, pcc_host_csr* out);
int pcc_write_adjacency_graph(const char* path, const pcc_host_csr* csr, int float_weighted);

/* ---------------------------------------------------------------- misc -- */

const char* pcc_last_error(void);
const char* pcc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PCC_AMD_H_ */
