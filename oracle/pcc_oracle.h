/*
 * pcc_oracle.h -- CPU restatement of the ParallelCorrelationClusterer /
 * ParallelModularityClusterer hot path (Louvain best moves, contraction,
 * refinement).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Determinism: single-threaded semantics; the
 * only OpenMP loop is the sync best-move map, whose per-vertex results depend
 * only on the round's snapshot, so thread count never changes a result.
 */
#ifndef PCC_ORACLE_H_
#define PCC_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_NO_MOVE 0xFFFFFFFFu

/* CorrelationClustererConfig fields on the path (clustering/config.proto:94-183). */
typedef struct oracle_config {
  double resolution;                  /* field 1 */
  double edge_weight_offset;          /* field 2 */
  int64_t num_iterations;             /* field 4 */
  int64_t louvain_num_iterations;     /* LouvainConfig field 1 */
  int64_t louvain_num_inner_iterations; /* LouvainConfig field 2 */
  int32_t clustering_moves_method;    /* 0 DEFAULT_CLUSTER_MOVES, 1 LOUVAIN */
  int32_t move_method;                /* 0 NBHR_CLUSTER_MOVE, 1 ALL_MOVE, 2 NBHR_MOVE */
  int32_t refine;
  int32_t async;
  int32_t permute;
  uint64_t permute_seed;              /* processing-order seed when permute */
  /* 1: the sequential CorrelationClusterer / ModularityClusterer instead
   * (clustering/clusterers/correlation-clusterer.cc, -internal.cc): vertices
   * move one after another in ascending order, no whole-cluster phase. */
  int32_t sequential;
  int32_t all_iter;                   /* field 17 (sequential clusterer only) */
  /* field 10: 0 NONE, 1 TRIANGLE, 2 CONNECTIVITY, 3 SLOW_CONNECTIVITY */
  int32_t subclustering_method;
} oracle_config;

/* Read-only CSR view: offsets[n+1], nbrs[m], weights[m] (float, as the
 * reference's symmetric_ptr_graph<symmetric_vertex, float>, IMC.h:52-53). */
typedef struct oracle_graph {
  int64_t n;
  int64_t m;
  const int64_t* offsets;
  const uint32_t* nbrs;
  const float* weights;
} oracle_graph;

/* ClusteringHelper state (PCCI.h:158-164), caller-owned, length n each. */
typedef struct oracle_state {
  uint32_t* cluster_ids;
  uint32_t* cluster_sizes;
  double* node_weights;
  double* cluster_weights;
} oracle_state;

void oracle_set_threads(int threads);
int oracle_get_threads(void);

/* Wall seconds spent by oracle_round / oracle_cluster since the last reset,
 * per phase: 0 best-move map, 1 sync apply, 2 frontier expansion,
 * 3 contraction, 4 refinement relabel.  reset != 0 zeroes them after copying. */
void oracle_phase_seconds(double out[5], int reset);

/* EfficientBestMove, PCCI.cc:448-583.  Returns (target, delta); target == n
 * means "form a new cluster". */
void oracle_best_move(const oracle_graph* g, const oracle_state* s,
                      const oracle_config* cfg, uint32_t v,
                      uint32_t* out_target, double* out_delta);

/* Every candidate EfficientBestMove evaluates for v (PCCI.cc:483-504 for
 * degree <= 1000, best_move_seq of PCCI.cc:551-563 above): clusters[i] /
 * deltas[i] for i < min(count, cap) in evaluation order, and *out_change the
 * "new cluster" value (change_in_objective).  Returns the candidate count. */
int64_t oracle_best_move_candidates(const oracle_graph* g, const oracle_state* s,
                                    const oracle_config* cfg, uint32_t v, uint32_t* clusters,
                                    double* deltas, int64_t cap, double* out_change);

/* BestMove(graph, moving_nodes), PCCI.cc:644-707 with BestMoveFromStats,
 * UTIL.h:70-113: the best cluster for the whole set members[0..count) (in
 * the reference's visiting order).  target == n means "a new cluster". */
void oracle_best_move_set(const oracle_graph* g, const oracle_state* s,
                          const oracle_config* cfg, const uint32_t* members,
                          int64_t count, uint32_t* out_target, double* out_delta);

/* The sync vertex map of BestMovesForVertexSubset, PCC.cc:132-160: best move,
 * singleton rule, then moves[v] iff delta > 0.  frontier == NULL means all
 * vertices.  out_moves[v] = target or ORACLE_NO_MOVE; out_moved[v] in {0,1};
 * out_target/out_delta may be NULL.  Entries of vertices outside the frontier
 * are ORACLE_NO_MOVE / 0. */
void oracle_sync_decisions(const oracle_graph* g, const oracle_state* s,
                           const oracle_config* cfg, const uint8_t* frontier,
                           uint32_t* out_moves, uint8_t* out_moved,
                           uint32_t* out_target, double* out_delta);

/* MoveNodesToCluster(moves), PCCI.cc:256-369.  Updates s in place and writes
 * out_modified[n] (0/1). */
void oracle_apply_moves(int64_t n, oracle_state* s, const uint32_t* moves,
                        uint8_t* out_modified);

/* One BestMovesForVertexSubset round, PCC.cc:115-338 (sync or async per cfg).
 * frontier (n bytes) is read and overwritten with the next frontier.  Returns
 * the size of the next frontier. */
int64_t oracle_round(const oracle_graph* g, oracle_state* s,
                     const oracle_config* cfg, uint8_t* frontier,
                     uint64_t round_seed);

/* IterateBestMoves, PCC.cc:340-365.  Returns 1 if "moved". */
int oracle_iterate_best_moves(const oracle_graph* g, oracle_state* s,
                              const oracle_config* cfg, int64_t inner,
                              uint64_t seed, int64_t* out_rounds);

/* CompressGraph, PCCI.cc:709-782 (+ ComputeInterClusterEdgesSort,
 * PGU.cc:156-231).  Returns an owned handle. */
typedef struct oracle_csr oracle_csr;
oracle_csr* oracle_compress(const oracle_graph* g, const uint32_t* cluster_ids,
                            const double* node_weights);
void oracle_csr_sizes(const oracle_csr* c, int64_t* n, int64_t* m);
void oracle_csr_copy(const oracle_csr* c, int64_t* offsets, uint32_t* nbrs,
                     float* weights, double* node_weights);
void oracle_csr_free(oracle_csr* c);

/* ClusteringHelper::ComputeObjective, PCCI.cc:118-149 (cluster weights are
 * recomputed from ids and node weights). */
double oracle_objective(const oracle_graph* g, const uint32_t* cluster_ids,
                        const double* node_weights, double resolution,
                        double edge_weight_offset);

/* ComputeModularity, PMC.cc:36-66. */
double oracle_modularity(const oracle_graph* g, const uint32_t* cluster_ids,
                         double resolution);

/* Whole pipeline: ParallelCorrelationClusterer::Cluster (PCC.cc:641-653) when
 * modularity == 0, ParallelModularityClusterer::Cluster (PMC.cc:98-111) when
 * modularity == 1.  node_weights == NULL means the default (1.0 for
 * correlation, degree for modularity).  initial_ids == NULL means singletons;
 * otherwise it is RefineClusters with that clustering (ids in [0, n)).
 * out_ids receives the final per-vertex cluster ids exactly as the reference's
 * cluster_ids vector before OutputIndicesById (PCC.cc:606-616).
 * clustering_moves_method 0 (DEFAULT_CLUSTER_MOVES) runs one level of
 * rounds that move vertices and then whole clusters (PCC.cc:167-297, 488-492).
 * Returns 0, or 12 (UNIMPLEMENTED) for an unsupported moves method. */
typedef struct oracle_result {
  double objective;          /* PCC.cc:619-624 */
  double modularity;         /* PMC.cc:91-93 (modularity runs only) */
  int64_t levels;            /* final iter + 1 */
  int64_t rounds;            /* best-move rounds, including refinement */
  int64_t edges_scanned;     /* sum over rounds of sum_{v in frontier} deg(v) */
  int64_t num_clusters;
  double seconds;            /* span PCC.cc:481 -> PCC.cc:609 */
} oracle_result;

int oracle_cluster(const oracle_graph* g, const oracle_config* cfg,
                   int modularity, const double* node_weights,
                   const uint32_t* initial_ids, uint32_t* out_ids,
                   oracle_result* out);

/* CompressSubclusters' numbering (PCC.cc:436-474, PCCI.cc:846-1120) of the
 * clustering ids: method 1 TRIANGLE, 2 CONNECTIVITY, 3 SLOW_CONNECTIVITY
 * (which reads edge_weight_offset and resolution).  sub[v] = the subcluster
 * id of v; parent[j] (length n) = rank of the cluster whose span holds id j,
 * for j below min(total, n).  Returns total = the sum of the spans, or -1
 * for another method. */
int64_t oracle_subclusters(const oracle_graph* g, const uint32_t* ids, int method,
                           uint32_t* sub, uint32_t* parent, double offset, double res);

/* CountTriangles' per-edge score, PCCI.cc:786-799. */
double oracle_triangle_score(const oracle_graph* g, uint32_t u, uint32_t v);

/* FlattenClustering, PGU.cc:233-243. */
void oracle_flatten(int64_t n, const uint32_t* ids, const uint32_t* compressed,
                    uint32_t* out);

/* Compact rank relabel as OutputIndicesById + SetClustering produce it
 * (PSO.h:183-209, PCCI.cc:60-77): id -> rank of id among distinct ids. */
void oracle_rank_relabel(int64_t n, const uint32_t* ids, uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif  /* PCC_ORACLE_H_ */
