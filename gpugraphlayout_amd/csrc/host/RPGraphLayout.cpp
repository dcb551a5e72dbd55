#include "RPGraphLayout.hpp"

#include <zlib.h>

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>

namespace RPGraph
{
    GraphLayout::GraphLayout(UGraph& graph, float width, float height)
        : graph(graph), width(width), height(height), xy(2 * (size_t)graph.num_nodes(), 0.0f)
    {
    }

    void GraphLayout::randomizePositions()
    {
        for (nid_t i = 0; i < graph.num_nodes(); ++i)
        {
            setX(i, get_random(-width / 2.0, width / 2.0));
            setY(i, get_random(-height / 2.0, height / 2.0));
        }
    }

    float GraphLayout::minX()
    {
        float v = std::numeric_limits<float>::max();
        for (nid_t n = 0; n < graph.num_nodes(); ++n) v = std::fmin(v, getX(n));
        return v;
    }
    float GraphLayout::maxX()
    {
        float v = -std::numeric_limits<float>::max();
        for (nid_t n = 0; n < graph.num_nodes(); ++n) v = std::fmax(v, getX(n));
        return v;
    }
    float GraphLayout::minY()
    {
        float v = std::numeric_limits<float>::max();
        for (nid_t n = 0; n < graph.num_nodes(); ++n) v = std::fmin(v, getY(n));
        return v;
    }
    float GraphLayout::maxY()
    {
        float v = -std::numeric_limits<float>::max();
        for (nid_t n = 0; n < graph.num_nodes(); ++n) v = std::fmax(v, getY(n));
        return v;
    }
    float GraphLayout::getXRange() { return maxX() - minX(); }
    float GraphLayout::getYRange() { return maxY() - minY(); }
    float GraphLayout::getSpan() { return std::ceil(std::fmax(getXRange(), getYRange())); }

    float GraphLayout::getDistance(nid_t n1, nid_t n2)
    {
        const float dx = getX(n1) - getX(n2), dy = getY(n1) - getY(n2);
        return std::sqrt(dx * dx + dy * dy);
    }

    Real2DVector GraphLayout::getDistanceVector(nid_t n1, nid_t n2)
    {
        return Real2DVector(getX(n2) - getX(n1), getY(n2) - getY(n1));
    }

    Coordinate GraphLayout::getCenter()
    {
        return Coordinate(minX() + getXRange() / 2.0, minY() + getYRange() / 2.0);
    }

    // CSV: "id,x,y\n" per node, id as in the edge list, default stream formatting
    // (src/RPGraphLayout.cpp:209-226).
    bool GraphLayout::writeToCSV(std::string path)
    {
        if (is_file_exists(path))
        {
            printf("Error: File exists at %s\n", path.c_str());
            return false;
        }
        std::ofstream out(path);
        for (nid_t n = 0; n < graph.num_nodes(); ++n)
            out << graph.node_map_r[n] << "," << getX(n) << "," << getY(n) << "\n";
        return out.good();
    }

    // BIN: packed {uint32 id, float x, float y} per node in internal order
    // (src/RPGraphLayout.cpp:228-250).
    bool GraphLayout::writeToBin(std::string path)
    {
        if (is_file_exists(path))
        {
            printf("Error: File exists at %s\n", path.c_str());
            return false;
        }
        std::ofstream out(path, std::ofstream::binary);
        for (nid_t n = 0; n < graph.num_nodes(); ++n)
        {
            const nid_t id = graph.node_map_r[n];
            const float x = getX(n), y = getY(n);
            out.write(reinterpret_cast<const char*>(&id), sizeof(id));
            out.write(reinterpret_cast<const char*>(&x), sizeof(x));
            out.write(reinterpret_cast<const char*>(&y), sizeof(y));
        }
        return out.good();
    }

    namespace
    {
        // Per-pixel hit counters for nodes and edges; y up, as the reference's canvas.
        struct Counts
        {
            int w, h;
            std::vector<uint32_t> nodes, edges;
            Counts(int w, int h) : w(w), h(h), nodes((size_t)w * h, 0), edges((size_t)w * h, 0) {}
            static void hit(std::vector<uint32_t>& c, int w, int h, long x, long y)
            {
                if (x < 0 || y < 0 || x >= w || y >= h) return;
                c[(size_t)(h - 1 - y) * w + x] += 1;
            }
            void disc(float cx, float cy)
            {
                const long x0 = std::lround(cx), y0 = std::lround(cy);
                for (int dy = -3; dy <= 3; ++dy)
                    for (int dx = -3; dx <= 3; ++dx)
                        if (dx * dx + dy * dy <= 9) hit(nodes, w, h, x0 + dx, y0 + dy);
            }
            // ceil(max(|dx|, |dy|)) + 1 evenly spaced samples (positions formed in double),
            // each rounded to a pixel.
            void line(float x0, float y0, float x1, float y1)
            {
                const float ddx = x1 - x0, ddy = y1 - y0;
                const float len = std::fmax(std::fabs(ddx), std::fabs(ddy));
                const int steps = (int)std::ceil(len);
                if (steps == 0)
                {
                    hit(edges, w, h, std::lround(x0), std::lround(y0));
                    return;
                }
                for (int k = 0; k <= steps; ++k)
                {
                    const double t = (double)k / (double)steps;
                    const float sx = (float)((double)x0 + t * (double)ddx);
                    const float sy = (float)((double)y0 + t * (double)ddy);
                    hit(edges, w, h, std::lround(sx), std::lround(sy));
                }
            }
        };

        // b^k by repeated squaring (the device rasteriser uses the same sequence).
        float powi(float b, uint32_t k)
        {
            float r = 1.0f;
            while (k)
            {
                if (k & 1u) r = r * b;
                b = b * b;
                k >>= 1;
            }
            return r;
        }

        void put32(std::vector<unsigned char>& v, uint32_t x)
        {
            v.push_back(x >> 24);
            v.push_back(x >> 16);
            v.push_back(x >> 8);
            v.push_back(x);
        }

        void chunk(FILE* fh, const char* type, const std::vector<unsigned char>& data)
        {
            std::vector<unsigned char> head;
            put32(head, (uint32_t)data.size());
            head.insert(head.end(), type, type + 4);
            fwrite(head.data(), 1, head.size(), fh);
            if (!data.empty()) fwrite(data.data(), 1, data.size(), fh);
            uLong crc = crc32(0L, (const Bytef*)type, 4);
            if (!data.empty()) crc = crc32(crc, data.data(), (uInt)data.size());
            std::vector<unsigned char> tail;
            put32(tail, (uint32_t)crc);
            fwrite(tail.data(), 1, 4, fh);
        }
    }

    gfa2_raster png_frame(float minx, float maxx, float miny, float maxy, int image_w, int image_h, nid_t nodes,
                          nid_t edges)
    {
        gfa2_raster f;
        const float xr = maxx - minx, yr = maxy - miny;
        const float cx = minx + xr / 2.0, cy = miny + yr / 2.0;
        f.x0 = cx - xr / 2.0;
        f.y0 = cy - yr / 2.0;
        f.sx = xr > 0 ? image_w / xr : 1.0f;
        f.sy = yr > 0 ? image_h / yr : 1.0f;
        f.node_alpha = std::fmin(1.0f, (float)(10000.0 / std::max<nid_t>(nodes, 1)));
        f.edge_alpha = std::fmin(1.0f, (float)(100000.0 / std::max<nid_t>(edges, 1)));
        f.width = image_w;
        f.height = image_h;
        return f;
    }

    void render_gray(const gfa2_raster& f, const UGraph& graph, const float* xy, std::vector<unsigned char>& gray)
    {
        Counts cv(f.width, f.height);
        const nid_t V = graph.num_nodes();
        for (nid_t n1 = 0; n1 < V; ++n1)
        {
            const float ax = (xy[2 * n1] - f.x0) * f.sx, ay = (xy[2 * n1 + 1] - f.y0) * f.sy;
            cv.disc(ax, ay);
            for (nid_t n2 : graph.neighbors_with_geq_id(n1))
                cv.line(ax, ay, (xy[2 * n2] - f.x0) * f.sx, (xy[2 * n2 + 1] - f.y0) * f.sy);
        }
        const float keep_n = 1.0f - f.node_alpha, keep_e = 1.0f - f.edge_alpha;
        gray.resize((size_t)f.width * f.height);
        for (size_t p = 0; p < gray.size(); ++p)
            gray[p] = (unsigned char)std::lround(255.0f * (powi(keep_n, cv.nodes[p]) * powi(keep_e, cv.edges[p])));
    }

    // PNG of the layout: the layout's bounding box fitted to the image, white background,
    // nodes as discs of radius 3 with opacity 10000/V, edges as lines with opacity
    // 100000/E, both black and blended (src/RPGraphLayout.cpp:170-207; the reference draws
    // through the pngwriter library, which is not part of its checkout, so the pixels
    // themselves are this build's own rasterisation).
    bool GraphLayout::writeToPNG(const int image_w, const int image_h, std::string path)
    {
        if (image_w <= 0 || image_h <= 0) return false;
        const gfa2_raster f = png_frame(minX(), maxX(), minY(), maxY(), image_w, image_h, graph.num_nodes(),
                                        graph.num_edges());
        std::vector<unsigned char> gray;
        render_gray(f, graph, xy.data(), gray);
        return write_png_gray(path, image_w, image_h, gray.data());
    }

    bool write_png_gray(const std::string& path, int image_w, int image_h, const unsigned char* gray)
    {
        std::vector<unsigned char> raw;
        raw.reserve((size_t)image_h * (3 * image_w + 1));
        for (int y = 0; y < image_h; ++y)
        {
            raw.push_back(0);
            for (int x = 0; x < image_w; ++x)
            {
                const unsigned char v = gray[(size_t)y * image_w + x];
                raw.push_back(v);
                raw.push_back(v);
                raw.push_back(v);
            }
        }
        uLongf zlen = compressBound(raw.size());
        std::vector<unsigned char> z(zlen);
        if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) return false;
        z.resize(zlen);
        FILE* fh = fopen(path.c_str(), "wb");
        if (!fh) return false;
        const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        fwrite(sig, 1, 8, fh);
        std::vector<unsigned char> ihdr;
        put32(ihdr, (uint32_t)image_w);
        put32(ihdr, (uint32_t)image_h);
        ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});   // 8-bit RGB, deflate, no filter, no interlace
        chunk(fh, "IHDR", ihdr);
        chunk(fh, "IDAT", z);
        chunk(fh, "IEND", {});
        return fclose(fh) == 0;
    }
}
