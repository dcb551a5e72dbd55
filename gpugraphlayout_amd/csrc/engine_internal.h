// Internal declarations shared by the engine's HIP translation units.
//
// Device data layout of one layout instance (V bodies, E undirected edges):
//   pos      float2[V]   positions, caller's node order
//   mass     float[V]    deg+1 unless given
//   rowptr   int[V+1], col int[2E]   symmetric CSR
//   fga      float2[V]   gravity + attraction of the current iteration
//   f        float2[V]   total force of the current iteration
//   fprev    float2[V]   total force of the previous iteration (F')
//   keys     u64[V] x2, vals u32[V] x2   quadtree path keys and body ids (sort ping-pong)
//   lcp      i8[V]       common-prefix length (levels) of sorted neighbours i, i+1
//   S        int[V+1]    exclusive prefix of cells opened at each sorted position
//   rec      float4[V+C] the quadtree in pre-order: {x, y, mass, info}
//            info == -1: a body; otherwise a cell, (skip << 5) | (level - 1) as unsigned,
//            skip being the number of records in its subtree (the walk jumps by it)
//   brec     int[V]      record index of the body at sorted position k
//   rcount   int[V+C]    body count of each cell record
// Tree levels: the root is level 0; a key holds 32 levels of 2-bit quadrant digits, so
// cells exist down to level 32. Bodies that share all 32 digits share one level-32 cell.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

struct gfa2_engine;

namespace gfa2 {

constexpr int kKeyLevels = 32;          // quadrant digits per key
constexpr int kMaxLevels = kKeyLevels + 1;
constexpr int kInfoLevelBits = 5;
constexpr int kMaxRecords = (1 << 27) - 2;  // skip must fit in 27 bits of info
constexpr int kPartBlocks = 1024;       // bounding-box partials (displacement grid cap)

// Error flags raised on the device, read back at synchronisation points.
constexpr int kErrTreeOverflow = 1;
constexpr int kErrSortWait = 2;         // a sort look-back wait ran out (results invalid)

// Per-instance scalar state, resident in device memory.
struct DevState {
    float rootx, rooty, radius, pad0;
    float global_speed, speed_efficiency, jitter_tolerance, k_s_max;
    double swing, traction;
    unsigned long long interactions;
    int ncells;
    int nrecords;
    int max_level;      // deepest cell level
    int err;
    int iteration;
    int nbbox;          // number of valid bounding-box partials
    int nbig;           // heavy cells listed for the in-order summary
    int sort_skip;      // low 8-bit digit passes the next path-key sort skips (0: a full sort)
    int full_sorts;     // iterations after new positions whose next sort is still a full one
    int runmax[8];      // largest body count of a cell at level 4, 8, ..., 32 (the last tree)
    int level_count[kMaxLevels + 1];
    int level_offset[kMaxLevels + 2];
    int level_cursor[kMaxLevels + 1];
};

// Prefix sums over sorted bodies for the integer-mass summary path: mass, and mass times the
// position relative to the graph's root centre in fixed point (coordinate * 2^s, with s per
// graph from its root radius so every coordinate is below 2^62 in magnitude). All sums are
// integers modulo 2^128 (two 64-bit limbs), so a cell's moment, the difference of two prefix
// sums, is exact: it does not depend on the bodies before the cell, on the other graphs of a
// batch or on how the sums were tiled.
struct Moment {
    unsigned long long xl, xh;   // sum m * qx, low and high limb
    unsigned long long yl, yh;   // sum m * qy
    unsigned long long m;        // sum m
};
constexpr int kMaxBig = 2048;   // cells too heavy for an exact prefix-sum mass (>= 2^24)

struct WalkPart {
    float swing, traction;
    unsigned int interactions, pad;
};

struct SortBuffers {
    unsigned long long* keys[2];
    unsigned int* vals[2];
    unsigned int* hist;      // 256 x tiles (three-launch passes)
    unsigned int* rowtot;    // 256
    bool onesweep = true;    // one launch per pass with decoupled look-back
    unsigned int* digit_hist = nullptr;   // 8 x 256: digit totals of every pass
    bool hist_in_keys = false;            // k_keys counts the digits; the fix-up clears the table
    unsigned int* status = nullptr;       // tiles x 256: published tile counts / prefixes
    unsigned long long* ticket = nullptr; // tile tickets, never reset while the tile count holds
    int* err = nullptr;                   // the instance's error word: kErrSortWait is raised there if a
                                          // look-back wait runs out (never expected)
};

// FA2 scalar parameters (src/RPForceAtlas2.cpp:28-55 defaults).
struct Params {
    int strong_gravity;
    float k_g, k_r;
    float itolsq, epssq;
    int group;              // walk vote group: 32 or 64 lanes
    int exact = 0;          // 1: exact O(V^2) repulsion instead of the tree and walk
    int linlog = 0;         // attraction log(1+d)/d per unit vector
    int walk_bt = 256;      // walk threads per block (64..512); results do not depend on it
};

// Launch shapes the autotuner chooses from; none of them changes any result bit.
struct LaunchConfig {
    int walk_bt;            // 64, 128, 256, 512
    int blocks_per_cu;      // grid of the grid-stride kernels (keys, displacement, bbox)
    int attraction_stream;  // 0: one stream; 1: attraction on a second stream beside a split walk
                            // (k_combine then forms F); 2: second stream beside keys, sort, tree
    int attraction_bpc;     // attraction grid per CU (0: one block per virtual block)
};

constexpr int kAttrChunk = 4096;    // entries per block for rows over 2,048 entries

// All device buffers of one instance.
struct Bufs {
    int n;                  // V
    long long nnz;          // 2E
    int grid_cap = 2048;    // block cap of the grid-stride kernels (blocks_per_cu * CUs)
    int cap_records;        // capacity of rec / rcount (V + max cells)
    float2* pos;
    float* mass;
    int* rowptr;
    int* col;
    int* node_lists;        // light | medium | heavy node ids
    int nlight, nmedium, nheavy;
    // Heavy rows are split into chunks of kAttrChunk entries, one block each; a second kernel
    // adds each heavy row's chunk partials in chunk order.
    int nattr_chunks;
    int attr_grid = 0;      // k_attract launch blocks (0: one per virtual block)
    int2* attr_chunks;      // {heavy index h, first entry}
    int* attr_c0;           // [nheavy + 1] first chunk of each heavy row
    float2* attr_part;      // [nattr_chunks]
    // Attraction gathers neighbour positions from a copy ordered by decreasing degree: the
    // positions most edges point at share few cache lines. pos_rank[rank[v]] = pos[v] and
    // col_rank[e] = rank[col[e]], so pos_rank[col_rank[e]] is pos[col[e]], bit for bit.
    int* rank;
    int* col_rank;
    float2* pos_rank;
    float2* fga;
    float2* f;
    float2* fprev;
    float2* rep;            // repulsion capture (debug), may be null
    SortBuffers sort;
    signed char* lcp;
    int* blocksum;
    int* S;
    float4* rec;
    int4* kids;             // per cell record: the record indices of its 1-4 children (-1: none)
    int* kids_part;         // [kids_blocks][8]: per k_kids block, the largest level-4p cell
    int kids_blocks = 1;
    int* brec;
    float4* sbody;          // {x, y, mass, 0} of the body at sorted position k (node k in exact mode)
    float4* pm;             // {x, y, mass, 0} of node v, written with the keys (k_lcp gathers it)
    float* ew_raw = nullptr;    // edge weights per CSR entry (null: unweighted)
    float* ew = nullptr;        // w^delta per CSR entry
    int* rcount;
    int* cell_rec;
    unsigned char* cell_lvl;
    int* level_list;
    float4* bbox_part;
    WalkPart* walk_part;
    int nwalk_part;
    DevState* st;
    // Batched layout of many graphs (nseg > 1). Graph g owns nodes [node_off[g],
    // node_off[g+1]) of the concatenated input and, after the sort, the same range of sorted
    // positions; its records are [rec_begin[g], rec_begin[g+1]).
    int nseg = 1;
    int seg_bits = 0;           // bits of a graph id
    const int* seg_of = nullptr;      // node -> graph
    int* node_off = nullptr;          // nseg + 1
    int* sg = nullptr;                // sorted position -> graph
    unsigned long long* pkey = nullptr;   // path key per node (regathered after the graph sort)
    float4* seg_root = nullptr;       // {rootx, rooty, radius, 0} per graph
    int* rec_begin = nullptr;         // nseg + 1
    int2* groups = nullptr;           // walk groups {first sorted position, graph}, per graph in order
    int ngroups = 0;
    int* grp_first = nullptr;         // graph -> first group, nseg + 1
    int* seg_order = nullptr;         // graphs by decreasing node count (per-graph sort)
    int max_seg = 0;                  // largest graph's node count
    struct SegState* seg_st = nullptr;
    WalkPart* grp_part = nullptr;     // per group
    // Integer masses (all whole numbers, total < 2^24 * 62): a cell lighter than 2^24 has
    // the same float mass in every summation order, so it is summarised from prefix sums;
    // only heavier cells are summed in quadrant order.
    bool int_mass = false;
    bool no_big = false;              // total mass below 2^24: no cell is summed in quadrant order
    Moment* pmass = nullptr;          // n + 1
    Moment* bmass = nullptr;          // per tree tile
    int* chunk_sum = nullptr;         // per 256 tree tiles (the two-level tile scan)
    Moment* chunk_mass = nullptr;
    int* big_rc = nullptr;            // kMaxBig
    unsigned char* big_lvl = nullptr; // kMaxBig
};

// Per-graph FA2 state of a batch.
struct SegState {
    float global_speed, speed_efficiency, jitter_tolerance, k_s_max;
    double swing, traction;
    unsigned long long interactions;
    int iteration, pad;
    unsigned long long interactions_total;   // summed over all iterations since reset
};

// Batched-mode launch groups (batch.hip).
void launch_seg_bbox(const Bufs& b, hipStream_t s);
void launch_seg_keys(const Bufs& b, hipStream_t s);
void launch_seg_graph_sort(Bufs& b, hipStream_t s);     // after the path sort

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kTreeThreads = 256;
constexpr int kTreeItems = 2;
constexpr int kTreeTile = kTreeThreads * kTreeItems;
// up to this many tiles k_prefix sums the tile totals before it itself (no scan launch)
constexpr int kFuseTiles = 2048;
constexpr int kTileChunk = 256;       // tree tiles per block of the two-level tile scan

// Radix sort of (key, value) pairs on 64-bit keys, ascending and stable. Sorts
// buf.keys[0]/vals[0] in place (via buf.keys[1]/vals[1]). With skip (path keys only, begin_bit
// 0), the passes below skip->sort_skip do nothing and the passes below `first` are not
// launched: the pairs end up sorted by the key's top bits only, possibly in buffer 1, and
// launch_sort_fixup (given the same `first`) then orders each run of equal top bits by the
// whole key (and value), which gives the full sort's order, in buffer 0.
// lead: the keys were written to the buffer from which the passes that run end in buffer 0
// (key_buffer); otherwise they start in buffer 0.
void radix_sort_pairs(SortBuffers& buf, int n, int begin_bit, int end_bit, hipStream_t s,
                      const DevState* skip = nullptr, int first = 0, bool lead = false, bool hist_ready = false);
// Whether radix_sort_pairs sorts n pairs with one launch per pass (the digit histograms of
// every pass taken first).
bool sort_onesweep(const SortBuffers& buf, int n);
// After radix_sort_pairs with a device skip: orders the runs of pairs that agree in every
// sorted bit by (whole key, value) and leaves the whole result in keys[0]/vals[0] (the passes
// that ran may have ended in the other buffer).
// seg: the graph of each sorted position when every graph was sorted in its own range (runs
// then stop at graph boundaries); null for one graph or a global sort.
// clear_hist: also clears buf.digit_hist for the next iteration's k_keys.
void launch_sort_fixup(SortBuffers& buf, int n, const DevState* st, hipStream_t s, int first = 0,
                       bool lead = false, const int* seg = nullptr, bool clear_hist = false);
// Passes of the path-key sort below which the fix-up finishes the order: the smallest p whose
// 4p-level cells of the previous tree hold at most this many bodies.
constexpr int kSortRunMax = 16;
// Sorts each segment [node_off[g], node_off[g+1]) of buf.keys[0]/vals[0] in place by the full
// 64-bit key, stable, one block per segment (taken in `order`); for batches whose graphs all
// have at most kSegSortMax nodes.
// skip: the state whose sort_skip low passes are skipped (then launch_sort_fixup with the
// graph of each position, and lead, finishes the order); null: all 8 passes.
void segmented_sort_pairs(SortBuffers& buf, const int* node_off, const int* order, int nseg, hipStream_t s,
                          const DevState* skip = nullptr);
constexpr int kSegSortMax = 32768;

// Launch groups of one iteration, in order.
void launch_attraction(const Bufs& b, const Params& p, hipStream_t s);
void launch_col_rank(const Bufs& b, hipStream_t s);
void launch_bbox_partials(const Bufs& b, hipStream_t s);
// first: the lowest path-key sort pass the host launches; the keys go to the buffer from which
// the passes that will run end in buffer 0.
void launch_keys(const Bufs& b, int first, hipStream_t s);
void launch_tree(Bufs& b, hipStream_t s);              // after the sort: cells + summaries
// split: the walk writes the repulsion alone (F_ga is not read), and launch_combine, after
// both the walk and the attraction, forms F and the swing/traction sums.
void launch_walk(const Bufs& b, const Params& p, bool split, hipStream_t s);
void launch_combine(const Bufs& b, const Params& p, hipStream_t s);
void launch_speed(const Bufs& b, hipStream_t s);
void launch_displacement(const Bufs& b, hipStream_t s);
// Exact mode: pack {pos, mass} in node order, then all-pairs repulsion per graph with the
// walk's epilogue (total force, swing/traction partials per group).
void launch_exact(const Bufs& b, const Params& p, hipStream_t s);
// All-pairs repulsion of the current positions into out (double2 per node), FP32 or FP64 sums.
void launch_exact_only(const Bufs& b, const Params& p, bool fp64, double2* out, hipStream_t s);
// w^delta of every CSR entry.
void launch_edge_weights(const Bufs& b, float delta, hipStream_t s);

// Layout quality (stress.hip): hop distances of K*T sampled pairs by multi-source BFS,
// and the normalized-stress sums {sum r, sum r^2, pairs} for positions pos.
int stress_hop_distances(const Bufs& b, const int* d_src, int K, const int* d_tgt, int T, int* d_dist,
                         unsigned long long* seen, unsigned long long* front, unsigned long long* next, int* d_flag,
                         hipStream_t s);
int stress_reduce(const float2* pos, const int* d_src, const int* d_tgt, const int* d_dist, int T, long long np,
                  double* d_part, int nblocks, double out[3], hipStream_t s);

// Access to an engine's device positions for code that writes them on the device (the
// learned initialiser): the view, and the notification that they were written.
struct EngineView {
    int n;
    int device;
    float2* pos;
    hipStream_t stream;
    const int* rowptr;      // the graph's device CSR
    const int* col;
    long long nnz;
    int nseg;
};
bool engine_view(struct ::gfa2_engine* h, EngineView* v);
// The buffers, device and stream of an engine holding a graph and positions (rasteriser).
bool engine_graph(struct ::gfa2_engine* h, const Bufs** b, int* device, hipStream_t* stream);
int engine_positions_written(struct ::gfa2_engine* h);

// Walk grid sizes (partials count depends on them).
int walk_blocks(int ngroups, int group, int bt);
int displacement_blocks(const Bufs& b);

}  // namespace gfa2
