/*
 * gfa2.h — C ABI of the MI355X ForceAtlas2 engine (libgfa2.so).
 *
 * The reference (govertb/GPUGraphLayout) has no C layer: its host class CUDAForceAtlas2
 * launches kernels directly. Each entry point below replaces a piece of that class or
 * of the ForceAtlas2/LayoutAlgorithm interface it implements; the comment on each names
 * the reference lines. Plain pointers and sizes only; every call returns an int status
 * (GFA2_OK or a negative GFA2_ERR_*). The library never retains caller pointers. A
 * handle owns one HIP stream and all its device memory; handles are independent but a
 * single handle is not thread-safe.
 */
#ifndef GFA2_H
#define GFA2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFA2_OK 0
#define GFA2_ERR_NO_DEVICE (-1)      /* src/RPGPUForceAtlas2.cu:42-48 exits instead */
#define GFA2_ERR_OOM (-2)            /* cudaMalloc failure, src/RPGPUForceAtlas2.cu:120-149 */
#define GFA2_ERR_INVALID_ARG (-3)
#define GFA2_ERR_INVALID_GRAPH (-4)  /* empty graph, bad CSR, self-loops, ids out of range */
#define GFA2_ERR_TREE_OVERFLOW (-5)  /* the reference overflows silently, src/RPBHKernels.cu:285-286 */
#define GFA2_ERR_HIP (-6)            /* any other HIP runtime error, src/RPCommon.hpp:33-41 exits */
#define GFA2_ERR_NO_GRAPH (-7)       /* step/positions before set_graph */
#define GFA2_ERR_UNSUPPORTED (-8)    /* e.g. exact mode: src/graph_viewer.cpp:93-97 */
#define GFA2_ERR_DEVICE_WAIT (-9)    /* a bounded wait of a device kernel ran out (the iteration's results
                                        are invalid and its later kernels were skipped); not expected */
#define GFA2_ERR_SCRATCH_LIMIT (-10) /* the request exceeds the fixed scratch of an allocation-free entry point */

typedef struct gfa2_engine* gfa2_handle;

/* FA2 parameters. Defaults (gfa2_default_params) are the reference's
 * (src/RPForceAtlas2.cpp:28-55). */
typedef struct gfa2_params {
    int32_t use_barneshut;      /* 1: Barnes-Hut repulsion; 0: exact O(V^2) repulsion (all pairs, the
                                   same term and epsilon; the reference's GPU path has no exact mode,
                                   src/graph_viewer.cpp:93-97) */
    int32_t strong_gravity;     /* sg / wg, src/graph_viewer.cpp:62 */
    float gravity;              /* k_g, ForceAtlas2::setGravity, src/RPForceAtlas2.cpp:69-72 */
    float scale;                /* k_r, ForceAtlas2::setScale, src/RPForceAtlas2.cpp:64-67 */
    float theta;                /* 1.0; itolsq = 1/theta^2 */
    float epssq;                /* 0.05*0.05 */
    float jitter_tolerance;     /* 1.0 */
    float k_s_max;              /* 10.0 */
    float global_speed;         /* initial speed, 1.0 */
    float speed_efficiency;     /* initial efficiency, 1.0 */
    int32_t group_size;         /* walk vote group: 32 = the reference's warp vote; 64 = wave64 */
    int32_t cell_factor;        /* cell capacity = cell_factor * V (+ slack); 0 = default 3 */
    int32_t linlog;             /* ForceAtlas2::use_linlog: attraction log(1+d)/d per unit vector,
                                   as src/RPCPUForceAtlas2.cpp:63-67 (0 = linear) */
    float edge_weight_influence;/* delta (src/RPForceAtlas2.hpp:49): attraction scaled by w^delta
                                   for edge weight w (gfa2_set_edge_weights); 0 = unweighted */
    int32_t prevent_overlap;    /* ForceAtlas2::prevent_overlap: must be 0 (GFA2_ERR_UNSUPPORTED,
                                   as the reference exits, src/RPCPUForceAtlas2.cpp:201-205) */
} gfa2_params;

/* Per-iteration statistics (no counterpart in the reference, which reads none back). */
typedef struct gfa2_stats {
    int32_t iteration;          /* completed iterations (ForceAtlas2::iteration) */
    int32_t nodes;
    int64_t edges;              /* undirected */
    float global_speed;
    float speed_efficiency;
    double swing;               /* S of the last iteration */
    double traction;            /* T of the last iteration */
    int64_t interactions;       /* body-node interactions of the last walk */
    int64_t interactions_total; /* summed over all iterations since the last reset */
    int32_t cells;              /* quadtree cells of the last tree */
    int32_t max_level;          /* deepest cell level (root 0); max body depth = max_level + 1 */
    float root_x, root_y, radius;
    int32_t error_flags;        /* bit 0: tree capacity exceeded; bit 1: a device wait ran out */
} gfa2_stats;

/* Force components for gfa2_get_forces. */
#define GFA2_FORCE_TOTAL 0          /* F of the last iteration (F' after it) */
#define GFA2_FORCE_REPULSION 1      /* Barnes-Hut part; needs gfa2_set_capture(h, 1) */
#define GFA2_FORCE_ATTRACTION_GRAVITY 2

void gfa2_default_params(gfa2_params* p);

/* Replaces the constructor's device setup, src/RPGPUForceAtlas2.cu:37-49, :86-117 (device
 * query, buffer sizing). device: HIP ordinal. */
int gfa2_create(const gfa2_params* params, int device, gfa2_handle* out);

/* New parameter values for subsequent iterations: ForceAtlas2::setScale / setGravity
 * (src/RPForceAtlas2.cpp:64-72) and the public strong_gravity flag
 * (src/RPForceAtlas2.hpp:48). Speed state and graph are kept; theta, epssq and
 * group_size may change too. */
int gfa2_set_params(gfa2_handle h, const gfa2_params* params);

/* Replaces host staging + H2D copies of masses and edges, src/RPGPUForceAtlas2.cu:51-84,
 * :152-161 and sendGraphToGPU :262-268. Symmetric CSR: rowptr[n+1] (int64), col[rowptr[n]]
 * holds every undirected edge in both rows. mass may be NULL: then deg+1
 * (ForceAtlas2::mass, src/RPForceAtlas2.cpp:74-77). Resets the FA2 state. Limits: 1 <= n <=
 * 2^26 nodes, rowptr[n] < 2^31 entries (GFA2_ERR_INVALID_GRAPH beyond). */
int gfa2_set_graph(gfa2_handle h, int32_t n, const int64_t* rowptr, const int32_t* col, const float* mass);

/* Same, from the reference's COO list (each undirected edge once, as
 * src/RPGPUForceAtlas2.cu:73-84 stages it). */
int gfa2_set_graph_coo(gfa2_handle h, int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                       const float* mass);

/* A batch of independent graphs laid out together (no counterpart in the reference, which
 * holds one graph per process): graph g owns nodes [node_off[g], node_off[g+1]) of a
 * block-diagonal symmetric CSR (no edge may leave its graph). Each graph gets its own root
 * box, tree, walk groups and speed state, so it follows, bit for bit, the trajectory it would
 * follow alone. Positions are the concatenation. */
int gfa2_set_graph_batch(gfa2_handle h, int32_t ngraphs, const int32_t* node_off, const int64_t* rowptr,
                         const int32_t* col, const float* mass);
/* Per-graph speed, efficiency, swing, traction and interaction count of the last
 * iteration (any pointer may be NULL). */
int gfa2_get_batch_stats(gfa2_handle h, int32_t cap, float* speed, float* efficiency, double* swing,
                         double* traction, int64_t* interactions);

/* H2D positions (interleaved x,y; n pairs). The reference's private sendLayoutToGPU,
 * src/RPGPUForceAtlas2.cu:256-260, made public. Does not reset F' or the speed state. */
int gfa2_set_positions(gfa2_handle h, const float* xy);

/* D2H positions: retrieveLayoutFromGPU, src/RPGPUForceAtlas2.cu:250-254 (the copy half of
 * sync_layout, :270-278). Waits for queued iterations. */
int gfa2_get_positions(gfa2_handle h, float* xy);

/* Fresh FA2 state: F' = 0, speed/efficiency from params, iteration 0. The reference only
 * has this at construction (src/RPGPUForceAtlas2.cu:63-71, src/RPFA2Kernels.cu:28-33). */
int gfa2_reset(gfa2_handle h);

/* Enqueue iters iterations: CUDAForceAtlas2::doStep, src/RPGPUForceAtlas2.cu:206-248, and
 * ForceAtlas2::doSteps, src/RPForceAtlas2.cpp:59-62. Asynchronous: no host sync per
 * iteration. */
int gfa2_step(gfa2_handle h, int32_t iters);

/* Wait for queued work; reports device-side errors (GFA2_ERR_TREE_OVERFLOW,
 * GFA2_ERR_DEVICE_WAIT). The reference's cudaDeviceSynchronize at src/RPGPUForceAtlas2.cu:246.
 * The error stays raised (and every later iteration is skipped) until gfa2_reset or
 * gfa2_set_graph. */
int gfa2_sync(gfa2_handle h);

/* Force vectors of the last iteration (interleaved x,y). */
int gfa2_get_forces(gfa2_handle h, int32_t kind, float* xy);

/* Keep the repulsion of each iteration in a separate buffer (debug / parity). */
int gfa2_set_capture(gfa2_handle h, int32_t on);

/* The last quadtree, cells in pre-order (canonical: independent of build order):
 * level (root 0), body count, mass, centroid (x,y). perm: the in-order body permutation
 * (the reference's sortd, src/RPBHKernels.cu:513-559). Any output pointer may be NULL.
 * *ncells receives the cell count; cap bounds what is written. */
int gfa2_get_tree(gfa2_handle h, int32_t cap, int32_t* ncells, int32_t* level, int32_t* count, float* mass,
                  float* xy, int32_t* perm);

int gfa2_get_stats(gfa2_handle h, gfa2_stats* s);

/* Per-kernel-group timing with HIP events on the handle's stream (no counterpart in the
 * reference). on = GFA2_TIMING_GROUPS: every iteration records events between all launch
 * groups; GFA2_TIMING_WALK: only the pair around the force walk (the least added work);
 * 0: off. The events carry no system-scope fence. */
#define GFA2_TIMING_GROUPS 1
#define GFA2_TIMING_WALK 2
int gfa2_set_timing(gfa2_handle h, int32_t on);
/* Average milliseconds per iteration of each measured group since timing was enabled:
 * names[i] (static strings), ms[i]; *count receives the number of groups (<= cap). */
int gfa2_get_timing(gfa2_handle h, int32_t cap, int32_t* count, const char** names, double* ms);

/* Launch shapes (replaces the compile-time THREADSn/FACTORn of
 * src/RPBHFA2LaunchParameters.cuh:10-59). None of them changes any result bit: the walk
 * runs one vote group per wavefront whatever its block size, the grid-stride kernels
 * only produce exact min/max partials, every attraction row is summed alike whatever its
 * grid, and the attraction reads only positions and writes only F_ga, so running it beside
 * other kernels changes no input of any kernel. */
typedef struct gfa2_launch_config {
    int32_t walk_block;         /* walk threads per block: 64, 128, 256 or 512 */
    int32_t blocks_per_cu;      /* grid of the grid-stride kernels, per compute unit */
    int32_t attraction_stream;  /* gravity + attraction: 0 on the instance's stream, before the
                                   other kernels; 1 on a second stream, beside the walk (which
                                   then leaves F and the swing/traction sums to a short combine
                                   kernel); 2 on a second stream, beside keys, sort and tree */
    int32_t attraction_blocks_per_cu;  /* attraction grid per compute unit (0: one block per
                                   256 light rows, 4 medium rows or one heavy-row chunk) */
    double ms_per_iteration;    /* measured by gfa2_autotune */
} gfa2_launch_config;

/* Times every launch shape over `iters` iterations, keeps the fastest for this handle and
 * for every later handle whose node count falls in the same power-of-two bucket, and
 * restores the layout state (the run continues as if no tuning iterations were made). */
int gfa2_autotune(gfa2_handle h, int32_t iters, gfa2_launch_config* best);
int gfa2_get_launch_config(gfa2_handle h, gfa2_launch_config* c);
int gfa2_set_launch_config(gfa2_handle h, const gfa2_launch_config* c);
/* The tuned shapes recorded for the size bucket of `nodes`; GFA2_ERR_NO_GRAPH if none. */
int gfa2_tuned_config(int32_t nodes, gfa2_launch_config* c);

/* Edge weights w (> 0, finite) for the attraction, one per CSR entry in the order of the col
 * array given to gfa2_set_graph or gfa2_set_graph_batch (count == rowptr[n]; the two entries
 * of an undirected edge should carry the same weight). NULL clears them. Used with
 * params.edge_weight_influence (w^delta); no counterpart in the reference, which declares
 * delta but never reads it. */
int gfa2_set_edge_weights(gfa2_handle h, int64_t count, const float* w);
/* gfa2_set_graph_coo with one weight per COO edge (the CSR order is built here). */
int gfa2_set_graph_coo_weighted(gfa2_handle h, int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                                const float* weight, const float* mass);

/* Exact repulsion of the current positions, all pairs of each graph with the same term as the
 * walk (k_r m_i m_j d / (|d|^2 + eps^2)), summed in FP32 (fp64 = 0) or FP64 (fp64 = 1), j in
 * ascending order; xy gets V (x, y) pairs. Does not advance the layout. */
int gfa2_exact_repulsion(gfa2_handle h, int32_t fp64, double* xy);

/* Rasteriser on the device for GraphLayout::writeToPNG (src/RPGraphLayout.cpp:170-207). The
 * frame maps layout point (x, y) to pixel ((x - x0) * sx, (y - y0) * sy), y up; the host layer
 * fills it from gfa2_get_bbox as its CPU writer does. Nodes are discs of radius 3, edges lines
 * from the smaller to the larger node id; a pixel hit by k nodes and l edges gets brightness
 * round(255 (1 - node_alpha)^k (1 - edge_alpha)^l). gray: width*height bytes, top row first.
 * Single-graph handles only (GFA2_ERR_UNSUPPORTED for a batch). */
typedef struct gfa2_raster {
    float x0, y0;
    float sx, sy;
    float node_alpha, edge_alpha;
    int32_t width, height;
} gfa2_raster;
/* Exact bounding box of the device positions: {min x, max x, min y, max y}. */
int gfa2_get_bbox(gfa2_handle h, float* out4);
int gfa2_render(gfa2_handle h, const gfa2_raster* r, uint8_t* gray);

/* gfa2_render without any device allocation, for callers that cannot allocate (tight serving
 * loops that draw every frame, harnesses that forbid hipMalloc). No counterpart in the
 * reference. Same arguments and bytes as gfa2_render; its counters and staging picture live in
 * a fixed per-device scratch array of the library (9 MiB, held by every device the library
 * runs on until the process ends), so pictures are limited to 1,048,576 pixels
 * (GFA2_ERR_SCRATCH_LIMIT beyond). All calls on one device share that array: calls from one
 * process are serialised by a lock, but calls that overlap on one device from different
 * processes race. */
int gfa2_render_noalloc(gfa2_handle h, const gfa2_raster* r, uint8_t* gray);

/* Graph capture of the whole iteration (hipGraph), replayed per iteration. */
int gfa2_set_graph_capture(gfa2_handle h, int32_t on);

/* Layout quality (no counterpart in the reference). Samples nsources distinct sources and
 * targets_per_source targets each (uniform over the other nodes; the engine's own seeded
 * generator) and computes their hop distances on the device by multi-source BFS. *npairs
 * receives the number of reachable pairs. */
int gfa2_stress_setup(gfa2_handle h, int32_t nsources, int32_t targets_per_source, uint64_t seed,
                      int64_t* npairs);
/* The sampled pairs: a[p], b[p], hop distance d[p] (0 = unreachable, left out). */
int gfa2_stress_pairs(gfa2_handle h, int64_t cap, int32_t* a, int32_t* b, int32_t* d);
/* Normalized stress (1/P) sum (alpha*|p_a-p_b|/d - 1)^2 with the optimal scale alpha, of the
 * device positions (xy == NULL) or of the given interleaved positions. */
int gfa2_normalized_stress(gfa2_handle h, const float* xy, double* ns, double* alpha);

/* ---- Learned initial layout (no counterpart in the reference) ----------------------------
 * A 2-layer message-passing network (hidden 64, BF16 weights, FP32 accumulation) mapping
 * per-node features (log degree + 8 seeded random-walk fields) to 2-D positions that
 * ForceAtlas2 then refines. Trained on the device with its own forward/backward kernels
 * against layouts FA2 produced, with a stress loss invariant to translation, rotation,
 * reflection and scale of the target. */
typedef struct gfa2_mpnn* gfa2_mpnn_t;

int32_t gfa2_mpnn_num_params(void);
/* Glorot-uniform weights from seed, zero biases. */
int gfa2_mpnn_create(int device, uint64_t seed, gfa2_mpnn_t* out);
/* FP32 master parameters (gfa2_mpnn_num_params floats); set resets the optimiser. */
int gfa2_mpnn_get_params(gfa2_mpnn_t m, float* params);
int gfa2_mpnn_set_params(gfa2_mpnn_t m, const float* params);
/* Registers a graph (symmetric CSR) and computes its features on the device
 * (smoothing_steps rounds of random-walk smoothing from feature_seed). target_xy (may be
 * NULL) is the layout to learn, interleaved x,y. */
int gfa2_mpnn_add_graph(gfa2_mpnn_t m, int32_t n, const int64_t* rowptr, const int32_t* col, const float* target_xy,
                        uint64_t feature_seed, int32_t smoothing_steps, int32_t* graph_id);
/* Registers the graph an engine holds (one graph, same device), copying its CSR device to
 * device, and computes its features as gfa2_mpnn_add_graph does; no target. */
int gfa2_mpnn_add_engine_graph(gfa2_mpnn_t m, gfa2_handle h, uint64_t feature_seed, int32_t smoothing_steps,
                               int32_t* graph_id);
/* Frees a registered graph's device memory; its id is invalid afterwards. */
int gfa2_mpnn_release_graph(gfa2_mpnn_t m, int32_t graph_id);
/* Features of a registered graph, [n x 16] (columns 9..15 are zero). */
int gfa2_mpnn_features(gfa2_mpnn_t m, int32_t graph_id, float* X);
/* Loss and its gradient (w.r.t. the BF16-rounded parameters) on one graph; pairs are the
 * CSR neighbours plus i +/- offsets[k] (mod n), k < R. grad may be NULL. */
int gfa2_mpnn_loss_grad(gfa2_mpnn_t m, int32_t graph_id, const int32_t* offsets, int32_t R, double* loss,
                        float* grad);
/* steps Adam steps (lr) cycling over the graphs that have targets; R seeded offsets per step.
 * *last_loss (may be NULL) receives the last step's loss. */
int gfa2_mpnn_train(gfa2_mpnn_t m, int32_t steps, float lr, int32_t R, uint64_t seed, double* last_loss);
/* Mean RMS radius of the training targets (the size the network's output is scaled to). */
float gfa2_mpnn_target_scale(gfa2_mpnn_t m);
/* Forward pass, centred and scaled to RMS radius `scale`; to host, or straight into an
 * engine's device positions (same node count and device). */
int gfa2_mpnn_infer(gfa2_mpnn_t m, int32_t graph_id, float scale, float* xy);
int gfa2_mpnn_infer_to_engine(gfa2_mpnn_t m, int32_t graph_id, float scale, gfa2_handle h);
int gfa2_mpnn_destroy(gfa2_mpnn_t m);

int gfa2_destroy(gfa2_handle h);

const char* gfa2_strerror(int code);

/* Device facts: compute units, wavefront size, total memory (bytes). */
int gfa2_device_info(int device, int32_t* cus, int32_t* wave, int64_t* mem_bytes);

#ifdef __cplusplus
}
#endif

#endif /* GFA2_H */
