#include "RPGraph.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <unordered_set>

namespace RPGraph
{
    namespace
    {
        // Parses "s t" (whitespace separated unsigned ids) at the start of a line.
        bool parse_pair(const char* p, nid_t& s, nid_t& t)
        {
            char* end = nullptr;
            while (*p == ' ' || *p == '\t') ++p;
            if (*p < '0' || *p > '9') return false;
            unsigned long a = strtoul(p, &end, 10);
            p = end;
            while (*p == ' ' || *p == '\t' || *p == ',') ++p;
            if (*p < '0' || *p > '9') return false;
            unsigned long b = strtoul(p, &end, 10);
            s = (nid_t)a;
            t = (nid_t)b;
            return true;
        }

        const char kMagic[8] = {'G', 'F', 'A', '2', 'C', 'S', 'R', '1'};
    }

    UGraph::UGraph(std::string path)
    {
        if (path.size() > 7 && path.compare(path.size() - 7, 7, ".csrbin") == 0)
        {
            loaded = load_binary(path);
            return;
        }
        FILE* fh = fopen(path.c_str(), "r");
        if (!fh) return;
        std::vector<nid_t> s, t;
        char* line = nullptr;
        size_t cap = 0;
        ssize_t len;
        while ((len = getline(&line, &cap, fh)) != -1)
        {
            if (len == 0 || line[0] == '#') continue;   // comments (src/RPGraph.cpp:46)
            nid_t a, b;
            if (!parse_pair(line, a, b)) continue;
            s.push_back(a);
            t.push_back(b);
        }
        free(line);
        fclose(fh);
        build(s, t);
        loaded = true;
    }

    UGraph::UGraph(const std::vector<nid_t>& s, const std::vector<nid_t>& t)
    {
        build(s, t);
        loaded = true;
    }

    void UGraph::build(const std::vector<nid_t>& s, const std::vector<nid_t>& t)
    {
        std::unordered_set<uint64_t> seen;
        seen.reserve(s.size() * 2);
        node_map.reserve(s.size());
        for (size_t k = 0; k < s.size(); ++k)
        {
            const nid_t a = s[k], b = t[k];
            if (a == b) continue;                                  // self-loop
            const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            if (!seen.insert(key).second) continue;                // repeated edge
            nid_t ma, mb;
            auto ia = node_map.find(a);
            if (ia == node_map.end())
            {
                ma = node_count++;
                node_map.emplace(a, ma);
                node_map_r.push_back(a);
            }
            else ma = ia->second;
            auto ib = node_map.find(b);
            if (ib == node_map.end())
            {
                mb = node_count++;
                node_map.emplace(b, mb);
                node_map_r.push_back(b);
            }
            else mb = ib->second;
            if (adjacency.size() < node_count)
            {
                adjacency.resize(node_count);
                degrees.resize(node_count, 0);
            }
            adjacency[ma < mb ? ma : mb].push_back(ma < mb ? mb : ma);
            degrees[ma] += 1;
            degrees[mb] += 1;
            ++edge_count;
        }
    }

    void UGraph::coo(std::vector<int32_t>& src, std::vector<int32_t>& dst) const
    {
        src.clear();
        dst.clear();
        src.reserve(edge_count);
        dst.reserve(edge_count);
        for (nid_t v = 0; v < node_count; ++v)
            for (nid_t u : adjacency[v])
            {
                src.push_back((int32_t)v);
                dst.push_back((int32_t)u);
            }
    }

    bool UGraph::save_binary(const std::string& path) const
    {
        FILE* fh = fopen(path.c_str(), "wb");
        if (!fh) return false;
        std::vector<int32_t> s, d;
        coo(s, d);
        uint64_t n = node_count, m = edge_count;
        bool ok = fwrite(kMagic, 1, 8, fh) == 8 && fwrite(&n, 8, 1, fh) == 1 && fwrite(&m, 8, 1, fh) == 1 &&
                  fwrite(node_map_r.data(), sizeof(nid_t), n, fh) == n && fwrite(s.data(), 4, m, fh) == m &&
                  fwrite(d.data(), 4, m, fh) == m;
        fclose(fh);
        return ok;
    }

    bool UGraph::load_binary(const std::string& path)
    {
        FILE* fh = fopen(path.c_str(), "rb");
        if (!fh) return false;
        char magic[8];
        uint64_t n = 0, m = 0;
        bool ok = fread(magic, 1, 8, fh) == 8 && memcmp(magic, kMagic, 8) == 0 && fread(&n, 8, 1, fh) == 1 &&
                  fread(&m, 8, 1, fh) == 1 && n < (1ull << 32) && m < (1ull << 32);
        std::vector<int32_t> s, d;
        if (ok)
        {
            node_map_r.resize(n);
            s.resize(m);
            d.resize(m);
            ok = fread(node_map_r.data(), sizeof(nid_t), n, fh) == n && fread(s.data(), 4, m, fh) == m &&
                 fread(d.data(), 4, m, fh) == m;
        }
        fclose(fh);
        if (!ok) return false;
        node_count = (nid_t)n;
        edge_count = (nid_t)m;
        adjacency.assign(n, {});
        degrees.assign(n, 0);
        for (nid_t v = 0; v < node_count; ++v) node_map.emplace(node_map_r[v], v);
        for (uint64_t e = 0; e < m; ++e)
        {
            if (s[e] < 0 || d[e] < 0 || (uint64_t)s[e] >= n || (uint64_t)d[e] >= n || s[e] >= d[e]) return false;
            adjacency[s[e]].push_back((nid_t)d[e]);
            degrees[s[e]] += 1;
            degrees[d[e]] += 1;
        }
        return true;
    }
}
