// Undirected graph read from an edge list, with the reference's semantics
// (src/RPGraph.hpp:55-81, src/RPGraph.cpp:35-131): '#' lines skipped, self-loops and
// repeated undirected edges dropped, node ids numbered 0.. by first appearance (source
// before target), each edge kept once at its smaller internal id in file order. The
// reference checks duplicates by copying and scanning a neighbour vector per line
// (O(sum deg^2)); here a hash set of endpoint pairs makes loading O(E).
#ifndef RPGraph_hpp
#define RPGraph_hpp
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace RPGraph
{
    typedef uint32_t nid_t;   // node ids (src/RPGraph.hpp:34-36)
    typedef uint32_t eid_t;

    class UGraph
    {
    public:
        // Loads an edge list (text). A file ending in ".csrbin" is read as the binary
        // cache written by save_binary().
        explicit UGraph(std::string edgelist_path);
        // From an in-memory edge list (file ids, in file order).
        UGraph(const std::vector<nid_t>& s, const std::vector<nid_t>& t);
        UGraph() = default;

        nid_t num_nodes() const { return node_count; }
        nid_t num_edges() const { return edge_count; }
        nid_t degree(nid_t nid) const { return degrees[nid]; }
        nid_t in_degree(nid_t nid) const { return degree(nid); }
        nid_t out_degree(nid_t nid) const { return degree(nid); }
        // Neighbours with a larger internal id, in insertion order (returned by
        // reference, not copied as the reference does).
        const std::vector<nid_t>& neighbors_with_geq_id(nid_t nid) const { return adjacency[nid]; }

        std::unordered_map<nid_t, nid_t> node_map;   // file id -> internal id
        std::vector<nid_t> node_map_r;               // internal id -> file id

        // The COO list each undirected edge once (src < dst, by ascending src, insertion
        // order): the staging of src/RPGPUForceAtlas2.cu:73-84.
        void coo(std::vector<int32_t>& src, std::vector<int32_t>& dst) const;

        // Binary cache: "GFA2CSR1", n, m, file ids, COO. Reloading it skips parsing.
        bool save_binary(const std::string& path) const;
        bool ok() const { return loaded; }

    private:
        nid_t node_count = 0, edge_count = 0;
        std::vector<nid_t> degrees;
        std::vector<std::vector<nid_t>> adjacency;
        bool loaded = false;
        void build(const std::vector<nid_t>& s, const std::vector<nid_t>& t);
        bool load_binary(const std::string& path);
    };
}

#endif
