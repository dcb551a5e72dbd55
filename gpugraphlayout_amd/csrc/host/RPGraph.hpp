// Undirected graph read from an edge list, with the reference's semantics
// (src/RPGraph.hpp:55-81, src/RPGraph.cpp:35-131): '#' lines skipped, self-loops and
// repeated undirected edges dropped, node ids numbered 0.. by first appearance (source
// before target), each edge kept once at its smaller internal id in file order. The
// reference checks duplicates by copying and scanning a neighbour vector per line
// (O(sum deg^2)); here an open-addressing set of endpoint pairs makes loading O(E), the
// file is memory-mapped and parsed in place, and adjacency is one CSR array.
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

    // Neighbour list view (a range over the CSR adjacency).
    struct NeighborRange
    {
        const nid_t* b;
        const nid_t* e;
        const nid_t* begin() const { return b; }
        const nid_t* end() const { return e; }
        size_t size() const { return (size_t)(e - b); }
        nid_t operator[](size_t i) const { return b[i]; }
    };

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
        // Neighbours with a larger internal id, in insertion order (a view, not the copy
        // the reference returns).
        NeighborRange neighbors_with_geq_id(nid_t nid) const
        {
            return NeighborRange{adj.data() + adj_off[nid], adj.data() + adj_off[nid + 1]};
        }

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
        std::vector<uint64_t> adj_off;   // CSR of each edge at its smaller internal id
        std::vector<nid_t> adj;
        bool loaded = false;
        void build(const nid_t* s, const nid_t* t, size_t count);
        void finish(const std::vector<nid_t>& lo, const std::vector<nid_t>& hi);
        bool load_binary(const std::string& path);
    };
}

#endif
