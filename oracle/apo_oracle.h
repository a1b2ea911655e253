/*
 * apo_oracle.h -- CPU restatement of apo's Apollonius-diagram construction.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the
 * reported CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path (the HIP
 * library under apo_amd/) never links or calls it.
 *
 * It restates, in plain host C++17 with IEEE FP32/FP64 arithmetic and no FMA
 * contraction, the method of PlathC/apo:
 *   - Morton codes, stable sort, Karras hierarchy, refit  (apo/src/gpu/lbvh.cu:11-323)
 *   - kNN traversal + heap                                (apo/include/apo/gpu/lbvh.inl:8-105, heap.hpp:11-43)
 *   - quadrisector / trisectorVertex / tangent planes      (apo/include/apo/gpu/math.cuh:390-780)
 *   - updateCell / cellRadius / initial octahedron        (apo/include/apo/gpu/topology_update.inl:5-653)
 *   - getKnns, getCells, vertex/edge/bisector validation, output
 *                                                         (apo/include/apo/gpu/algorithm.inl:32-1307,1410-1646)
 * The reference's 32-lane warp order is replaced by a fixed 64-wide step
 * order with ascending lane order (the order the HIP kernels use).
 */
#ifndef APO_ORACLE_H
#define APO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference sample stream (app/common/samples.cpp:64-69) ---------- */
double orc_pcg_unit(uint32_t v);          /* ((w>>22)^w) / 4294967295 */
double orc_pcg_ldexp(uint32_t v);         /* ldexp((w>>22)^w, -32)  (apo/src/core/utils.cpp:6-12) */

/* ---- geometry (FP64), exported for property tests ------------------- */
/* s = 4 doubles each (x,y,z,r).  x1,x2 = 4 doubles.  Returns the mask /
 * count exactly as the reference functions do.                        */
uint32_t orc_quadrisector(const double* s1, const double* s2, const double* s3, const double* s4,
                          double* x1, double* x2);
uint32_t orc_trisector(const double* s1, const double* s2, const double* s3, double* x1, double* x2);
void     orc_tangent_planes(const double* s1, const double* s2, const double* s3, double* n1, double* n2);
void     orc_sphere_from_points(const double* p1, const double* p2, const double* p3, const double* p4,
                                double* out);

/* ---- Morton / LBVH -------------------------------------------------- */
/* Scene box of sphere extents in FP64, cast to float as the reference does
 * (apo/src/gpu/algorithm.cpp:5-25, algorithm.inl:1509).                */
void orc_scene_box(const float* xyzr, uint32_t n, float* box6 /* minx,miny,minz,maxx,maxy,maxz */,
                   double* minmax_radius /* 2 */);
uint64_t orc_morton64(uint32_t x, uint32_t y, uint32_t z);
/* The 6 joggled artificial sites (algorithm.inl:1437-1464), 6 x (x,y,z,r). */
void     orc_artificial_sites(const float* xyzr, uint32_t n, double* out24);
void     orc_morton_codes(const float* xyzr, uint32_t n, const float* box6, uint64_t* codes);

/* Full LBVH build.  Node layout as the reference: root 0, internal
 * [0,n-1), leaves [n-1,2n-1).  aabb = 6 floats per node (min,max).     */
int orc_lbvh_build(const float* xyzr, uint32_t n, uint64_t* sorted_codes, uint32_t* indices, int32_t* parent,
                   int32_t* left, int32_t* right, float* aabb, uint32_t* sizes, float* max_radius);

/* kNN over the BVH of sorted sites (sorted_xyzr = sites in Morton order).
 * One pass from scratch (iteration 0).  knn = n*k indices (sorted space),
 * knn_dist = n*k gaps.  Returns max traversal stack depth used.          */
int orc_knn(const float* sorted_xyzr, uint32_t n, const int32_t* left, const int32_t* right, const float* aabb,
            uint32_t k, uint32_t* knn, float* knn_dist);
/* Brute-force kNN with the same strict acceptance (tests only). */
void orc_knn_brute(const float* sorted_xyzr, uint32_t n, uint32_t k, uint32_t* knn, float* knn_dist);

/* ---- full construction ------------------------------------------------ */
typedef struct orc_config {
    uint32_t k;                 /* neighbours per kNN pass (reference default 16) */
    uint32_t vertex_capacity;   /* per-cell vertex records (152) */
    uint32_t edge_capacity;     /* per-cell edge records (264) */
    uint32_t new_edge_capacity; /* new-edge entries per update (128) */
    uint32_t bisector_capacity; /* bisectors per cell in bisector validation (512) */
    uint32_t max_iterations;    /* vertex+edge validation rounds (10) */
    uint32_t max_restarts;      /* restarts per cell per validation kernel (100) */
    uint32_t threads;           /* OpenMP threads, 0 = default */
    uint32_t stage_mask;        /* bit0 knn+cells, bit1 vertex/edge validation, bit2 bisector validation */
    uint32_t edge_bound;        /* edge searches pruned by the cover of the edge's empty spheres: 0 never
                                   (the reference's search), 1 when the search sphere exceeds 2x the known
                                   radius, 2 always; results are identical in every mode */
    uint32_t fallback_k;        /* learned proposal: neighbours of the last class (not certified within 64), 16 */
    uint32_t knn_passes;        /* kNN + cell passes, the reference's knnValidationStep (algorithm.cuh:16,
                                   algorithm.inl:1535-1549): pass s > 0 continues each list after the previous
                                   pass's last neighbour; 0 runs none (cells stay the octahedron).  Default 1 */
    uint32_t complete;          /* completion mode (no counterpart in the reference), 0 = off (default): the
                                   hyperbolic edge searches prune a subtree only when no point of its box lies
                                   inside both tangent-plane half-spaces; the cells the bisector stage changed
                                   go through vertex and edge validation again; then up to `complete`
                                   consistency rounds: a cell holding fewer records of a quadruple (i,j,k,l)
                                   than one of its other three cells is updated with the smallest quadruple
                                   site it does not know, and the cells that changed are validated again */
    uint32_t lanes;             /* order of the validation loops: 64 (default) steps of 64 records with
                                   contributors applied lowest lane first, the HIP kernels' order; 32 the
                                   reference's 32-lane warps visited highest lane first (utils.cuh:17-32) */
    uint32_t bisector_bound;    /* bisector searches skip subtrees whose box cannot hold a site passing the
                                   cylinder test (1, default, as the HIP kernels) or search as the reference
                                   (0); results are identical either way */
} orc_config;

void orc_default_config(orc_config* cfg);

typedef struct orc_diagram orc_diagram;

orc_diagram* orc_build(const float* xyzr, uint32_t n, const orc_config* cfg);
/* The learned-proposal path given per-site classes (input order): 64-nearest
 * lists, each cell runs the neighbour count of its class (20, 24, 28, 32,
 * 40, 48, 64; class 7: cfg->fallback_k) before certification and validation. */
orc_diagram* orc_build_proposal(const float* xyzr, uint32_t n, const orc_config* cfg, const uint8_t* classes);
/* The range path given per-site search radii (input order): 64-long lists of
 * the sites whose gap is below max(radius, -2 r_i), nearest first; each cell
 * runs its whole list and, when the list is not full, takes the radius (less
 * the FP32 rounding of a gap) as its known radius. */
orc_diagram* orc_build_radii(const float* xyzr, uint32_t n, const orc_config* cfg, const float* radii);
void         orc_free(orc_diagram* d);

/* Emitted Apollonius vertices (apo toHost semantics). */
uint64_t orc_vertex_count(const orc_diagram* d);
void     orc_get_vertices(const orc_diagram* d, double* xyzw, uint32_t* ids);

/* Closed trisector edges (apo toHostFull semantics, algorithm.inl:1309-1383):
 * extreme point of smaller radius (x,y,z,w) and ids (i,j,k,0), original ids. */
uint64_t orc_closed_edge_count(const orc_diagram* d);
void     orc_get_closed_edges(const orc_diagram* d, double* xyzw, uint32_t* ids);

/* Per-cell state after the pipeline (sorted space). */
void orc_get_sorted_sites(const orc_diagram* d, float* xyzr, uint32_t* indices);
void orc_get_cell_counts(const orc_diagram* d, uint32_t* vertex_nb, uint32_t* edge_nb, uint8_t* status,
                         float* known_radius, uint32_t* flags);
/* Records of cell i: vertices as 4 u32 (j,k,l,type|status<<8), edges as 4 u32 (j,k,vertexNb,status). */
void orc_get_cell(const orc_diagram* d, uint32_t i, uint32_t* vertices, uint32_t* edges);
/* Concatenated records of all cells in cell order (counts from orc_get_cell_counts). */
void orc_get_all_cells(const orc_diagram* d, uint32_t* vertices, uint32_t* edges);
/* toTopology (algorithm.inl:1732-1821): counts3 = bisectors, trisectors,
 * quadrisectors; the arrays (2, 3, 4 u32 per entry, original ids) may be NULL. */
void orc_topology(const orc_diagram* d, uint64_t* counts3, uint32_t* bisectors, uint32_t* trisectors,
                  uint32_t* quadrisectors);
void orc_get_knns(const orc_diagram* d, uint32_t* knn);
/* Stage statistics: [0] uncertified after cells, [1] tasks per validation
 * iteration summed, [2] iterations run, [3] overflowed cells, [4] vertices
 * emitted, [5] buried cells, [6] cells changed by bisector validation,
 * [7] completion: consistency rounds run, [8] cells updated by them
 * (rebuilt ones included), [9] missing records whose sites the cell
 * already knows, counted by the last round that ran (left unrepaired),
 * [10] tasks of the validation rounds after the bisector stage and after
 * each consistency round, [11] cells rebuilt.                            */
void orc_get_stats(const orc_diagram* d, uint64_t* stats12);
/* Completion mode's subtree test: 1 when some point of the box (min xyz,
 * max xyz) may lie strictly inside the half-space n.(x - p) < 0. */
int orc_box_reaches_half_space(const float* box6, const float* n3, const float* p3);
/* Certified k per cell (first neighbour count at which the security-radius
 * test passes, 0xffffffff when the K neighbours never certify), computed
 * with the per-neighbour test of app/benchmarks/benchmark_contributing_nearest.cu:75-113. */
void orc_certified_k(const float* xyzr, uint32_t n, uint32_t kmax, uint32_t* out_k, uint32_t threads);

/* ---- checks ------------------------------------------------------------ */
/* Number of emitted vertices violated by some site (FP64 weighted distance
 * strictly below the vertex radius minus tol), checked by brute force.     */
uint64_t orc_empty_sphere_violations(const float* xyzr, uint32_t n, const double* xyzw, const uint32_t* ids,
                                     uint64_t v, double tol, uint32_t threads);
/* Interval precision validator (app/precision/precision.cpp:14-163,
 * apo_interval.cpp): Boost.Interval semantics with the rounding mode switched
 * per bound.  xyzw: count points (w unused); ids: count x 4 original ids
 * (arity 3: closed-edge points, ids[3] unused).  out2 = {bad, uncertain}. */
void orc_precision_counts(const float* xyzr, uint32_t n, const double* xyzw, const uint32_t* ids, uint64_t count,
                          uint32_t arity, uint32_t threads, uint64_t* out2);
/* One directed-rounding operation: op 0 a+b, 1 a*b, 2 sqrt(a); lohi = {down, up}. */
void orc_round_dir(uint32_t op, double a, double b, double* lohi);
/* Interval weighted distance |p - c| - r of one site (precision.cpp:14-24). */
void orc_interval_distance(const double* site4, const double* p3, double* lohi);

/* Cavity hit ratios (app/cavity/utils.cu:17-65, 201-242; apo_cavity.cpp):
 * the normalised Halton directions, and per vertex the fraction of the
 * rays from its FP32 centre that hit any site in [0.01, 1e16]. */
void orc_cavity_directions(uint32_t samples, float* xyz);
void orc_cavity_hit_ratios(const float* xyzr, uint32_t n, const double* xyzw, uint64_t count, uint32_t samples,
                           uint32_t threads, float* ratios);

/* All empty tangent spheres of 4 sites, by enumeration of quadruples
 * (tiny n only).  Returns the count; writes up to cap id quadruples sorted. */
uint64_t orc_bruteforce_vertices(const float* xyzr, uint32_t n, uint32_t* ids, uint64_t cap, double tol);

#ifdef __cplusplus
}
#endif

#endif /* APO_ORACLE_H */
