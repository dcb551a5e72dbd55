#include "RPGraph.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hashing.hpp"

namespace RPGraph
{
    namespace
    {
        inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

        // Unsigned decimal at p (a digit), saturating like strtoul; advances p.
        inline uint64_t parse_uint(const char*& p, const char* end)
        {
            uint64_t v = 0;
            bool sat = false;
            for (; p < end && is_digit(*p); ++p)
            {
                const uint64_t d = (uint64_t)(*p - '0');
                if (v > (UINT64_MAX - d) / 10) sat = true;
                else v = v * 10 + d;
            }
            return sat ? UINT64_MAX : v;
        }

        // Parses "s t" (ids separated by spaces, tabs or commas) at the start of the line
        // [p, end); anything after the second id is ignored.
        bool parse_pair(const char* p, const char* end, nid_t& s, nid_t& t)
        {
            while (p < end && (*p == ' ' || *p == '\t')) ++p;
            if (p == end || !is_digit(*p)) return false;
            const uint64_t a = parse_uint(p, end);
            while (p < end && (*p == ' ' || *p == '\t' || *p == ',')) ++p;
            if (p == end || !is_digit(*p)) return false;
            const uint64_t b = parse_uint(p, end);
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
        const int fd = open(path.c_str(), O_RDONLY);
        if (fd < 0) return;
        struct stat sb;
        if (fstat(fd, &sb) != 0)
        {
            close(fd);
            return;
        }
        const size_t size = (size_t)sb.st_size;
        std::vector<nid_t> s, t;
        if (size > 0)
        {
            void* map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (map == MAP_FAILED)
            {
                close(fd);
                return;
            }
            madvise(map, size, MADV_SEQUENTIAL);
            const char* p = static_cast<const char*>(map);
            const char* end = p + size;
            s.reserve(size / 16);
            t.reserve(size / 16);
            while (p < end)
            {
                const char* nl = static_cast<const char*>(memchr(p, '\n', (size_t)(end - p)));
                const char* le = nl ? nl : end;
                nid_t a, b;
                // comments (src/RPGraph.cpp:46)
                if (le > p && p[0] != '#' && parse_pair(p, le, a, b))
                {
                    s.push_back(a);
                    t.push_back(b);
                }
                p = nl ? nl + 1 : end;
            }
            munmap(map, size);
        }
        close(fd);
        build(s.data(), t.data(), s.size());
        loaded = true;
    }

    UGraph::UGraph(const std::vector<nid_t>& s, const std::vector<nid_t>& t)
    {
        build(s.data(), t.data(), s.size() < t.size() ? s.size() : t.size());
        loaded = true;
    }

    void UGraph::build(const nid_t* s, const nid_t* t, size_t count)
    {
        gfa2host::KeySet seen(count);
        gfa2host::IdMap ids(count / 4 + 1024);
        std::vector<nid_t> lo, hi;
        lo.reserve(count);
        hi.reserve(count);
        constexpr size_t kAhead = 16;
        for (size_t k = 0; k < count; ++k)
        {
            if (k + kAhead < count)
            {
                const nid_t pa = s[k + kAhead], pb = t[k + kAhead];
                seen.prefetch(pa < pb ? ((uint64_t)pa << 32 | pb) : ((uint64_t)pb << 32 | pa));
                ids.prefetch(pa);
                ids.prefetch(pb);
            }
            const nid_t a = s[k], b = t[k];
            if (a == b) continue;                                  // self-loop
            const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            if (!seen.insert(key)) continue;                       // repeated edge
            const uint32_t before = ids.size();
            const nid_t ma = ids.get_or_add(a);
            if (ids.size() != before) node_map_r.push_back(a);
            const uint32_t mid = ids.size();
            const nid_t mb = ids.get_or_add(b);
            if (ids.size() != mid) node_map_r.push_back(b);
            lo.push_back(ma < mb ? ma : mb);
            hi.push_back(ma < mb ? mb : ma);
        }
        node_count = (nid_t)node_map_r.size();
        node_map.reserve(node_count);
        for (nid_t v = 0; v < node_count; ++v) node_map.emplace(node_map_r[v], v);
        finish(lo, hi);
    }

    // CSR of the kept edges at their smaller id, in the order given (a stable counting sort).
    void UGraph::finish(const std::vector<nid_t>& lo, const std::vector<nid_t>& hi)
    {
        const size_t m = lo.size();
        edge_count = (nid_t)m;
        degrees.assign(node_count, 0);
        adj_off.assign((size_t)node_count + 1, 0);
        for (size_t e = 0; e < m; ++e)
        {
            degrees[lo[e]] += 1;
            degrees[hi[e]] += 1;
            adj_off[lo[e] + 1] += 1;
        }
        for (size_t v = 0; v < node_count; ++v) adj_off[v + 1] += adj_off[v];
        adj.resize(m);
        std::vector<uint64_t> cur(adj_off.begin(), adj_off.end() - 1);
        for (size_t e = 0; e < m; ++e) adj[cur[lo[e]]++] = hi[e];
    }

    void UGraph::coo(std::vector<int32_t>& src, std::vector<int32_t>& dst) const
    {
        src.resize(edge_count);
        dst.resize(edge_count);
        for (nid_t v = 0; v < node_count; ++v)
            for (uint64_t e = adj_off[v]; e < adj_off[v + 1]; ++e)
            {
                src[e] = (int32_t)v;
                dst[e] = (int32_t)adj[e];
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
        // The cached COO is ascending by source with src < dst; anything else is corrupt.
        for (uint64_t e = 0; e < m; ++e)
        {
            if (s[e] < 0 || d[e] < 0 || (uint64_t)s[e] >= n || (uint64_t)d[e] >= n || s[e] >= d[e]) return false;
            if (e > 0 && s[e] < s[e - 1]) return false;
        }
        node_count = (nid_t)n;
        node_map.reserve(n);
        for (nid_t v = 0; v < node_count; ++v) node_map.emplace(node_map_r[v], v);
        std::vector<nid_t> lo(s.begin(), s.end()), hi(d.begin(), d.end());
        finish(lo, hi);
        return true;
    }
}
