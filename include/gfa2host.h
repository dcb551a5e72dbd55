/*
 * gfa2host.h — C entry points of the host layer (libgfa2host.so): the edge-list loader and
 * the layout writers behind the C++ classes of RPGraph.hpp / RPGraphLayout.hpp, for
 * callers that are not C++ (bindings, tests). No device work.
 *
 *   gfa2host_load          UGraph(edgelist_path), src/RPGraph.cpp:35-56 (O(E) here)
 *   gfa2host_graph_arrays  node_map_r, degree(), and the COO staging of
 *                          src/RPGPUForceAtlas2.cu:73-84
 *   gfa2host_write_layout  writeToCSV / writeToBin / writeToPNG, src/RPGraphLayout.cpp:170-250
 */
#ifndef GFA2HOST_H
#define GFA2HOST_H
#include <stdint.h>
#include "gfa2.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gfa2host_graph* gfa2host_graph_t;

/* Loads an edge list (or a .csrbin cache). Returns 0, or -1 if unreadable. */
int gfa2host_load(const char* path, gfa2host_graph_t* out, int32_t* n, int64_t* m);
/* Builds a graph from an in-memory edge list in file order (file ids). */
int gfa2host_from_edges(int64_t count, const uint32_t* s, const uint32_t* t, gfa2host_graph_t* out, int32_t* n,
                        int64_t* m);
/* ids[n], deg[n], src[m], dst[m]; any may be NULL. */
void gfa2host_graph_arrays(gfa2host_graph_t g, int64_t* ids, int32_t* deg, int32_t* src, int32_t* dst);
/* Writes the binary cache (reloadable by gfa2host_load when named *.csrbin). */
int gfa2host_save_binary(gfa2host_graph_t g, const char* path);
/* fmt: "csv", "bin" or "png" (image_w x image_h). xy: interleaved positions. Refuses to
 * overwrite an existing file (returns -1). */
int gfa2host_write_layout(gfa2host_graph_t g, const float* xy, const char* path, const char* fmt, int32_t image_w,
                          int32_t image_h);
void gfa2host_free(gfa2host_graph_t g);

/* The PNG picture (src/RPGraphLayout.cpp:170-207): the frame fitting a bounding box to the
 * image, the CPU rasteriser (the same bytes gfa2_render produces on the device for the same
 * positions) and the PNG encoder for a width*height grey image (top row first). */
void gfa2host_png_frame(float minx, float maxx, float miny, float maxy, int32_t image_w, int32_t image_h,
                        int64_t nodes, int64_t edges, gfa2_raster* out);
int gfa2host_render_gray(gfa2host_graph_t g, const float* xy, const gfa2_raster* f, uint8_t* gray);
int gfa2host_write_png_gray(const char* path, int32_t image_w, int32_t image_h, const uint8_t* gray);

/* Barabasi-Albert preferential attachment, deterministic per seed (splitmix64): a star
 * K_{1,m} on nodes 0..m, then each node v = m+1..n-1 attaches to m distinct earlier nodes
 * drawn with probability proportional to degree (uniform draws from the endpoint list).
 * Writes m*(n-m-1)+m edges (v, u) in creation order, u ascending per v. Returns the count. */
int64_t gfa2host_barabasi_albert(int32_t n, int32_t m, uint64_t seed, int32_t* s, int32_t* t);

/* R-MAT edges over 2^scale ids (quadrant probabilities a, b, c, 1-a-b-c), drawn until
 * n_edges distinct undirected non-loop edges exist, in first-occurrence order with their
 * drawn orientation. Draw k, level l uses SplitMix64 output number k*scale+l of the stream
 * seeded with `seed`, as a double (z >> 11) * 2^-53; level 0 is the most significant id bit.
 * Returns the edge count (fewer than n_edges only if 64*n_edges+2^20 draws did not reach
 * it), or -1 for invalid arguments (scale outside 1..31, n_edges above half of all pairs). */
int64_t gfa2host_rmat(int32_t scale, int64_t n_edges, double a, double b, double c, uint64_t seed, int32_t* s,
                      int32_t* t);

#ifdef __cplusplus
}
#endif
#endif
