/*
 * apo_c.h -- C ABI of the MI355X (gfx950) Apollonius-diagram engine.
 *
 * Plain C types only: pointers, sizes and status codes.  One context owns
 * all device memory and one HIP stream; calls on one context are not
 * thread-safe, distinct contexts may be used from distinct threads.
 *
 * Each entry point names the reference interface it replaces
 * (PlathC/apo, paths relative to the repository root).
 */
#ifndef APO_C_H
#define APO_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum apo_status {
    APO_OK = 0,
    APO_ERR_HIP = 1,          /* a HIP runtime call failed (apo_last_error has the text) */
    APO_ERR_OUT_OF_MEMORY = 2,
    APO_ERR_INVALID = 3,      /* bad argument (null pointer, n < 2, unsupported k ...) */
    APO_ERR_STATE = 4,        /* call out of order (e.g. results before a build) */
    APO_ERR_CAPACITY = 5      /* caller buffer smaller than the reported count */
} apo_status;

typedef struct apo_ctx apo_ctx;

/* Construction parameters.  Defaults (apo_default_config) reproduce
 * AlgorithmGPU<16>(sites, knnValidationStep = 1)
 * (apo/include/apo/gpu/algorithm.cuh:12-16, algorithm.inl:1404-1594). */
typedef struct apo_config {
    uint32_t k;               /* neighbours per kNN pass, the template argument K of AlgorithmGPU<K>
                                 (algorithm.cuh:12): 1..64, default 16                 */
    uint32_t max_iterations;  /* vertex + edge validation rounds (10, algorithm.inl:1554) */
    uint32_t max_restarts;    /* restarts per cell per validation kernel (100, :272)  */
    uint32_t stage_mask;      /* bit0 cells, bit1 vertex/edge validation, bit2 bisector,
                                 bit3 stop after the kNN pass */
    uint32_t waves_per_cu;    /* waves per CU of the persistent cell and validation grids,
                                 0 = each kernel's full occupancy */
    uint32_t device;          /* HIP device ordinal */
    float proposal_cost;      /* learned proposal: cost of validating one cell, in neighbour updates;
                                 a site gets the class c minimising k_c + cost * P(not certified at k_c).
                                 0 = the most probable class */
    uint32_t graph;           /* 1: capture the whole device build (LBVH to vertex output) in a HIP graph
                                 and replay it while the site pointer, site and molecule counts and this
                                 config stay the same; stats then hold ms_total only */
    float cover_ratio;        /* edge validation: a search whose sphere exceeds cover_ratio x the cell's known
                                 radius is pruned by a cover of the edge's empty spheres (exact pruning: the
                                 results do not depend on it); < 0 = never.  Default 2 */
    uint32_t waves_cells;     /* waves per CU of each persistent grid (cell pass, vertex, edge and bisector */
    uint32_t waves_vertex;    /* validation); 0 = waves_per_cu when that is set, else full occupancy */
    uint32_t waves_edge;
    uint32_t waves_bisector;
    uint32_t proposal_fallback_k; /* learned proposal: neighbours for sites of the last class (not certified
                                     within 64): 16, 24, 32, 48 or 64; default 16 */
    uint32_t proposal_max_k;  /* learned proposal: the most neighbours any other class may run (16, 24, 32, 48
                                 or 64; default 64).  Classes asking for more are not chosen; the kNN list
                                 holds max(proposal_max_k, proposal_fallback_k) neighbours */
    uint32_t knn_passes;      /* kNN + cell passes: the knnValidationStep argument of AlgorithmGPU<K>
                                 (algorithm.cuh:16, algorithm.inl:1535-1549).  Pass s > 0 continues each
                                 site's list after the previous pass's k-th neighbour (a list the previous
                                 pass did not fill ends) and updates the cells still in validation; 0 runs
                                 no kNN or cell pass (every cell goes to validation as the octahedron).
                                 Default 1; the learned proposal needs 1 */
    uint32_t complete;        /* completion mode, no counterpart in the reference; 0 = off (default, the
                                 reference's output).  1..8: (a) the hyperbolic edge searches prune a subtree
                                 only when no point of its box lies inside both tangent-plane half-spaces (the
                                 reference tests the box point closest to the search sphere's centre,
                                 algorithm.inl:660-670, and skips contributors); (b) the cells the bisector
                                 stage changed go through vertex and edge validation again (the reference
                                 leaves them unvalidated, :1551-1594); (c) up to `complete` consistency rounds:
                                 a cell holding fewer records of a quadruple (i, j, k, l) than another of its
                                 four cells is updated with the smallest quadruple site it does not know (a
                                 cell that knows them all is rebuilt, its stuck records' sites first), and
                                 validated again.  The output then no longer depends on k or the learned
                                 proposal where the reference's does */
    uint32_t proposal_radius; /* learned proposal: 0 (default) each cell runs its class's neighbour count;
                                 1 each cell takes the sites below the MLP's search radius (the radius head)
                                 through a range query, as with apo_proposal_set_radii.  No counterpart in
                                 the reference */
    uint32_t cells_per_wave;  /* cell pass: cells a wave64 works on at once, 1 (the whole wave on one cell) or
                                 2 (half a wave each, twice the LDS per wave); 0 = default (1).  The records
                                 are the same either way; an autotuner knob */
    uint32_t bisector_per_cell; /* bisector validation: 0 (default) plans each cell's searches, runs them one per
                                   lane across cells and runs the per-cell stage only where a search finds a
                                   contributor; 1 runs every cell's stage per cell.  The records are the same */
} apo_config;

/* Per-build statistics (the stage split of AlgorithmGPU::Sample,
 * algorithm.cuh:18-27, in milliseconds, plus counts). */
typedef struct apo_stats {
    double ms_total;
    double ms_lbvh;
    double ms_knn;
    double ms_cells;
    double ms_validation;
    double ms_bisector;
    double ms_output;
    uint64_t sites;
    uint64_t vertices;          /* emitted Apollonius vertices */
    uint64_t uncertified;       /* cells not certified after the kNN pass */
    uint64_t buried;
    uint64_t flagged_cells;     /* cells with an overflow / cap flag */
    uint64_t vertex_records;    /* sum over cells of stored vertex records */
    uint64_t edge_records;      /* sum over cells of stored edge records */
    /* per-kernel device time (HIP events around each launch, summed) */
    double ms_vertex_validation_kernel;
    double ms_edge_validation_kernel;
    uint64_t validation_launches; /* rounds that launched the vertex + edge validation kernels */
    uint64_t validation_tasks;    /* cells processed, summed over those rounds */
    /* algorithmic HBM bytes of each kernel family (records, sites and per-cell
     * scalars each kernel must read or write once), summed over its launches */
    uint64_t bytes_cells;
    uint64_t bytes_vertex_validation;
    uint64_t bytes_edge_validation;
    uint64_t bytes_bisector_validation;
    /* learned proposal: features + MLP time, whether the build used it, and
     * the cells certified by the cell pass (before any validation) */
    double ms_proposal;
    uint64_t proposal;
    uint64_t cells_certified;
    /* completion mode (config.complete): its time, the consistency rounds that
     * ran, the cells they updated (rebuilt ones included), the records still
     * missing from a cell that knows their sites when the rounds end (counted
     * by the last round), the tasks of its validation rounds, and the cells
     * rebuilt because a missing record's sites were all known */
    double ms_completion;
    uint64_t completion_rounds;
    uint64_t completion_updated_cells;
    uint64_t completion_unrepairable;
    uint64_t completion_validation_tasks;
    uint64_t completion_rebuilt_cells;
    uint64_t reserved[1];
} apo_stats;

void apo_default_config(apo_config* cfg);

/* apo::joggle (apo/src/core/utils.cpp:14-25, core/utils.hpp:9): adds
 * pcg(4i + j) * 2 range - range to value j of site i, pcg(v) being the
 * shadertoy PCG hash scaled by 2^-32 (utils.cpp:6-12); values = 4n.  Host
 * only, no device needed.  The _f32 variant takes FP32 sites: the noise is
 * added in FP64 and each value rounded to FP32 once. */
void apo_joggle(double* xyzr, uint64_t values, double range);
void apo_joggle_f32(float* xyzr, uint64_t values, double range);

/* Replaces the AlgorithmGPU<K> constructor (algorithm.inl:1404-1408). */
apo_status apo_create(apo_ctx** ctx, const apo_config* cfg);
void apo_destroy(apo_ctx* ctx);
const char* apo_last_error(const apo_ctx* ctx);

/* Replaces AlgorithmGPU<K>::build() (algorithm.inl:1410-1594) including the
 * H2D copy of computeApolloniusDiagram (src/gpu/algorithm.cu:6-15).
 * xyzr: host array of n (x, y, z, r) FP32 quadruples, read during the call. */
apo_status apo_build(apo_ctx* ctx, const float* xyzr, uint32_t n);

/* Same, for sites already resident in device memory (n float4 records). */
apo_status apo_build_device(apo_ctx* ctx, const float* xyzr_device, uint32_t n);

/* ---- batches of independent molecules.  No counterpart in the reference,
 * whose callers build one molecule per AlgorithmGPU in a host loop
 * (app/benchmarks/full_benchmarks_proteins.cpp:43-79, one
 * computeApolloniusDiagram per protein).  Molecule m is sites
 * [offsets[m], offsets[m+1]) of xyzr (offsets: m+1 host values, offsets[0] = 0,
 * at least 2 sites per molecule, at most 2^20 molecules); its diagram is the
 * one apo_build gives for those sites alone, with site ids offset by
 * offsets[m].  All molecules go through one launch per stage: an
 * offset-indexed LBVH (one tree per molecule) and persistent work-queue
 * cell and validation kernels.  Vertices come out molecule after molecule;
 * artificial sites of molecule m have ids n + 6m + 0..5. ---- */
apo_status apo_build_batched(apo_ctx* ctx, const float* xyzr, const uint32_t* offsets, uint32_t molecules);
apo_status apo_build_batched_device(apo_ctx* ctx, const float* xyzr_device, const uint32_t* offsets,
                                    uint32_t molecules);
/* ---- a build that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * apo_build_device + apo_get_device_vertices for one molecule of 2 to
 * APO_STATIC_MAX_SITES sites on the plain path (any k and knn_passes of the
 * context's config), with every working buffer in one static device array of
 * APO_STATIC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends.  xyzr_device: n float4 sites in device memory;
 * xyzw_device / ids_device: device buffers for `capacity` vertices (FP64
 * x, y, z, w and 4 original site ids); *count receives the vertex count.  A
 * count above capacity returns APO_ERR_CAPACITY with the first `capacity`
 * vertices written.  Synchronous on the context's stream.  Calls that
 * overlap on one GPU (from any context) race on the array; after the call
 * the context holds no build (results are in the caller's buffers only). */
#define APO_STATIC_MAX_SITES 4096u
#define APO_STATIC_SCRATCH_BYTES (48u << 20)
apo_status apo_build_device_static(apo_ctx* ctx, const float* xyzr_device, uint32_t n, double* xyzw_device,
                                   uint32_t* ids_device, uint64_t capacity, uint64_t* count);

/* First output vertex of each molecule of the last build (molecules + 1 values). */
apo_status apo_batch_vertex_offsets(apo_ctx* ctx, uint64_t* offsets);

/* out5[0..3]: waves per CU at full occupancy of the persistent cell,
 * vertex, edge and bisector kernels on the context's device (the largest
 * useful waves_* settings); out5[4]: its CU count.  No counterpart in the
 * reference; read by the autotuner. */
apo_status apo_launch_occupancy(apo_ctx* ctx, uint32_t* out5);

/* Vertex output of AlgorithmGPU<K>::toHost (algorithm.inl:1596-1646):
 * each Apollonius vertex once, from the cell of its smallest site,
 * centre + empty-sphere radius in FP64, the 4 tangent sites as original
 * input indices.  Two-call pattern: query the count, then copy. */
apo_status apo_vertex_count(apo_ctx* ctx, uint64_t* count);
apo_status apo_get_vertices(apo_ctx* ctx, double* xyzw, uint32_t* ids, uint64_t capacity);
/* Device pointers to the vertex records (valid until the next build). */
apo_status apo_get_device_vertices(apo_ctx* ctx, const double** xyzw, const uint32_t** ids);

/* Closed trisector edges of AlgorithmGPU<K>::toHostFull (algorithm.inl:
 * 1648-1730, kernels getWriteableClosedEdges / saveClosedEdges at
 * :1309-1383; FullDiagram::closedEdgesMin / closedEdgesId of
 * apo/include/apo/gpu/algorithm.hpp:18-26): each closed elliptic trisector
 * once, from the cell of its smallest site, as the extreme point of smaller
 * empty-sphere radius (x, y, z, w in FP64) and ids (i, j, k, 0) in original
 * input indices, in cell order.  Two-call pattern: query the count, then
 * copy (either array may be NULL, not both). */
apo_status apo_closed_edge_count(apo_ctx* ctx, uint64_t* count);
apo_status apo_get_closed_edges(apo_ctx* ctx, double* xyzw, uint32_t* ids, uint64_t capacity);

apo_status apo_get_stats(apo_ctx* ctx, apo_stats* stats);

/* FP64 empty-sphere check through the LBVH of the last build, with the
 * criterion of app/precision/precision.cpp:26-100: a vertex is violated when
 * a site other than its 4 tangent sites lies closer than (its largest
 * tangent distance - tol).  xyzw/ids: host arrays of `count` vertices
 * (original site ids); pass NULL for both to check the build's own output. */
apo_status apo_check_empty_spheres(apo_ctx* ctx, const double* xyzw, const uint32_t* ids, uint64_t count, double tol,
                                   uint64_t* violations);

/* Interval precision validator of the reference's precision app
 * (app/precision/precision.cpp:14-163: getBadQuadrisectorNb and
 * getBadTrisectorNb), through the LBVH instead of a scan of every site.  For
 * each point, the weighted distances |p - c| - r to its own sites and to
 * every other site are FP64 intervals with outward rounding (the bounds of
 * Boost.Interval with directed rounding); a point is bad when some other
 * site's interval lies entirely below the largest of its own sites'
 * intervals, and uncertain (near-degenerate: FP64 cannot decide it) when it
 * is not bad and some other site's interval reaches below that one.  The
 * uncertain count has no counterpart in the reference. */
typedef struct apo_precision {
    uint64_t quadrisectors, bad_quadrisectors, uncertain_quadrisectors;  /* the build's vertices */
    uint64_t trisectors, bad_trisectors, uncertain_trisectors;           /* its closed edges */
} apo_precision;
/* The build's own vertices and closed edges. */
apo_status apo_precision_check(apo_ctx* ctx, apo_precision* out);
/* Host points: count x (x, y, z, w) (w unused) and count x 4 original site
 * ids; arity 4 for vertices, 3 for closed-edge points (ids[3] unused).
 * Each point is checked against the sites of its first site's molecule. */
apo_status apo_precision_check_points(apo_ctx* ctx, const double* xyzw, const uint32_t* ids, uint64_t count,
                                      uint32_t arity, uint64_t* bad, uint64_t* uncertain);

/* Cavity detection of the reference's cavity app (app/cavity/utils.cu;
 * the ray tracing there uses OptiX, here the LBVH).  Hit ratios
 * (getHitRatios, :17-65): for every vertex of the last build, the fraction of
 * `samples` rays (at most 4096) cast from its centre (as FP32) along Halton
 * directions on the sphere (sphereHalton, :201-242, normalised) that hit a
 * site sphere at a distance in [0.01, 1e16]; written to hit_ratios (one
 * float per output vertex, may be NULL) and kept for apo_cavity_vertices. */
apo_status apo_cavity_hit_ratios(apo_ctx* ctx, uint32_t samples, float* hit_ratios);
/* filterVertices (:244-264, :527-558): the vertices whose hit ratio is above
 * hit_ratio_threshold and whose radius (as FP32) is above radius_threshold,
 * in output order, as FP32 (x, y, z, w).  Two-call pattern: xyzw NULL
 * returns the count. */
apo_status apo_cavity_vertices(apo_ctx* ctx, float hit_ratio_threshold, float radius_threshold, float* xyzw,
                               uint64_t capacity, uint64_t* count);

/* Topology sets of AlgorithmGPU::toTopology (algorithm.inl:1732-1821): for
 * every cell in sorted order, its distinct bisectors {i, j}, trisectors
 * {i, j, k} and quadrisectors {i, j, k, l} with i below the others and no
 * artificial site, ascending within the cell, as original site ids (2, 3
 * and 4 u32 per entry).  counts3 receives the three entry counts; pass NULL
 * arrays to query them (two-call pattern). */
apo_status apo_get_topology_sets(apo_ctx* ctx, uint32_t* bisectors, uint64_t bisector_cap, uint32_t* trisectors,
                                 uint64_t trisector_cap, uint32_t* quadrisectors, uint64_t quadrisector_cap,
                                 uint64_t* counts3);

/* Per-site cell topology (the cell records of AlgorithmGPU's dVertices /
 * dEdges, algorithm.cuh:38-43) as CSR in original site order with original
 * ids: vertex records (j, k, l, type) and edge records (j, k, vertexNb, 0);
 * artificial sites keep ids n..n+5 (n + 6m.. in molecule m of a batch), vertexNb 0xffffffff marks a closed
 * edge, buried sites have no records.  vertex_offsets / edge_offsets: n+1
 * entries (may be NULL); NULL record arrays only fill the offsets. */
apo_status apo_get_cell_topology(apo_ctx* ctx, uint64_t* vertex_offsets, uint32_t* vertices, uint64_t vertex_cap,
                                 uint64_t* edge_offsets, uint32_t* edges, uint64_t edge_cap);

/* ---- learned proposal.  No counterpart in the reference's construction:
 * an MLP (52 -> 64 -> 64 -> 9, BF16 weights, FP32 accumulation on the
 * matrix cores) predicts per site how many nearest neighbours the cell pass
 * runs before the plain certification test (algorithm.inl:197-201), as 8
 * class logits, and the search radius its cell needs (output 8: log2 of the
 * radius over d16, the centre distance of the 16th nearest neighbour);
 * cells the test does not certify go through the plain validation.
 * Classes ask for 20, 24, 28, 32, 40, 48, 64 neighbours; class 7 (not
 * certified within 64) runs config.proposal_fallback_k (16 by default). ---- */
#define APO_PROPOSAL_FEATURES 52
#define APO_PROPOSAL_CLASSES 8
#define APO_PROPOSAL_OUTPUTS 9 /* the 8 class logits, then the radius head */
#define APO_PROPOSAL_PARAMS 8137 /* W1[52][64] b1[64] W2[64][64] b2[64] W3[64][9] b3[9], weights [in][out] */

/* Label pass (app/benchmarks/benchmark_contributing_nearest.cu:75-113): for
 * every site, the number of neighbours in kNN order after which the cell is
 * first certified, testing after every update (1..64; 0: not within 64;
 * 255: buried), plus its 52 features; both in input order. */
apo_status apo_proposal_labels(apo_ctx* ctx, const float* xyzr, uint32_t n, float* features, uint8_t* labels);
/* Train the MLP on (features, labels) from apo_proposal_labels with HIP
 * kernels on the matrix cores (the inference pass forward, BF16 weights and
 * activations with FP32 accumulation; an FP32 backward pass; Adam on FP32
 * master weights).  The loss: softmax cross-entropy over the 8 classes, plus
 * 0.25 x the squared error of the radius head against radius_targets (per
 * site, log2 of the wanted radius over d16; NaN or a NULL array: none).
 * init: starting parameters or NULL (seeded uniform); params receives
 * APO_PROPOSAL_PARAMS floats, epoch_loss (may be NULL) the mean loss of
 * each epoch. */
apo_status apo_proposal_train(apo_ctx* ctx, const float* features, const uint8_t* labels,
                              const float* radius_targets, uint64_t count, uint32_t epochs, uint32_t batch, float lr,
                              uint32_t seed, const float* init, float* params, float* epoch_loss);
/* Gradient of the training loss (above) averaged over `count` feature rows
 * with classes 0..7 and radius targets (may be NULL), at parameters
 * `params`, computed by the training kernels (BF16 forward pass as in
 * inference, FP32 backward pass on the matrix cores); grad receives
 * APO_PROPOSAL_PARAMS floats, loss (may be NULL) the mean loss.  For
 * checking the training kernels. */
apo_status apo_proposal_gradient(apo_ctx* ctx, const float* params, const float* features, const uint8_t* classes,
                                 const float* radius_targets, uint32_t count, float* grad, float* loss);
/* Use these parameters in the following builds (NULL: the plain path). */
apo_status apo_proposal_set(apo_ctx* ctx, const float* params);
/* Classes (and optionally the APO_PROPOSAL_OUTPUTS outputs: 8 logits and
 * the radius head) the set MLP gives to count feature rows. */
apo_status apo_proposal_infer(apo_ctx* ctx, const float* features, uint32_t count, uint8_t* classes, float* logits);
/* Search radii of the last build's sites (input order), when it ran the
 * range-query cell pass: the radius head's (config.proposal_radius) or the
 * ones given with apo_proposal_set_radii.  No counterpart in the reference. */
apo_status apo_proposal_radii(apo_ctx* ctx, float* radii);
/* Classes of the last build's sites, in input order. */
apo_status apo_proposal_classes(apo_ctx* ctx, uint8_t* classes);
/* Per-site classes (input order, values 0..7) for the following builds of n
 * sites, used instead of the MLP's: e.g. the classes of the labels of an
 * earlier build of the same sites (apo_proposal_labels).  Certification and
 * validation are unchanged.  NULL: back to the MLP (or the plain path). */
apo_status apo_proposal_set_classes(apo_ctx* ctx, const uint8_t* classes, uint32_t n);

/* Per-site search radii (input order) for the following builds of n sites,
 * used instead of neighbour counts: each cell is updated with every site
 * whose gap (|c_i - c_j| - r_i - r_j) is below its radius, nearest first, at
 * most max(proposal_max_k, proposal_fallback_k) of them, through a range
 * query over the LBVH; when the list holds fewer, every site below the
 * radius is known and the radius (less the FP32 rounding of a gap) serves as
 * the cell's known radius in the plain certification test
 * (algorithm.inl:197-201).  Cells the test does not certify go through the
 * plain validation.  A radius at or below -r_i is raised to the float just
 * above -r_i (sites whose spheres hold c_i stay in reach, the one that
 * buries site i among them).  Single builds only.  NULL: back to the MLP (or
 * the plain path).  No counterpart in the reference. */
apo_status apo_proposal_set_radii(apo_ctx* ctx, const float* radii, uint32_t n);

/* Per-site cell radius of the last build (input order): the largest
 * empty-sphere radius over the cell's vertices and the extreme points of its
 * closed trisector edges (cellRadius, topology_update.inl:595-653), the
 * quantity the certification test compares with half the known radius;
 * -FLT_MAX for buried sites.  Twice it is the smallest search radius that
 * certifies the final cell. */
apo_status apo_cell_radii(apo_ctx* ctx, float* radii);

/* ---- introspection of the device state (the public members dSites,
 * bvh, dStatus, dVertices, dEdges of AlgorithmGPU, algorithm.cuh:147-155) ---- */

/* Sorted sites (Morton order) and the sorted->original permutation. */
apo_status apo_get_sorted_sites(apo_ctx* ctx, float* xyzr, uint32_t* indices);
/* LBVH in the reference node numbering (lbvh.cuh:12-37): internal nodes
 * [0, n-1), leaves [n-1, 2n-1).  aabb: 6 floats per node (min, max);
 * max_radius: largest site radius below each node.  After a batch of m
 * molecules: 2n - m nodes, internal [0, n-m) (molecule i's tree rooted at
 * offsets[i] - i), leaves [n-m, 2n-m). */
apo_status apo_get_lbvh(apo_ctx* ctx, uint64_t* sorted_codes, int32_t* parent, int32_t* left, int32_t* right,
                        float* aabb, uint32_t* sizes, float* max_radius);
/* The 6 artificial bounding sites appended after the n input sites
 * (algorithm.inl:1437-1464), 6 x (x, y, z, r) in FP64. */
apo_status apo_get_artificial_sites(apo_ctx* ctx, double* xyzr24);
/* kNN lists of the first pass, n*k sorted-space indices. */
apo_status apo_get_knns(apo_ctx* ctx, uint32_t* knn);
/* Per-cell counts and status (sorted space). */
apo_status apo_get_cell_counts(apo_ctx* ctx, uint32_t* vertex_nb, uint32_t* edge_nb, uint8_t* status,
                               float* known_radius, uint32_t* flags);
/* Records of all cells concatenated in cell order: vertices as 4 u32
 * (j, k, l, type | status << 8), edges as 4 u32 (j, k, vertexNb, status). */
apo_status apo_get_all_cells(apo_ctx* ctx, uint32_t* vertices, uint32_t* edges);

#ifdef __cplusplus
}
#endif

#endif /* APO_C_H */
