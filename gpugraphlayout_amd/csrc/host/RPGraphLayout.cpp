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
        // Multiplicative ink buffer: a pixel covered with opacity a keeps (1-a) of its
        // brightness; products commute, so drawing order does not matter.
        struct Canvas
        {
            int w, h;
            std::vector<float> keep;
            Canvas(int w, int h) : w(w), h(h), keep((size_t)w * h, 1.0f) {}
            void blend(int x, int y, float a)
            {
                if (x < 0 || y < 0 || x >= w || y >= h) return;
                keep[(size_t)(h - 1 - y) * w + x] *= (1.0f - a);   // y up, as the reference's canvas
            }
            void disc(float cx, float cy, int r, float a)
            {
                const int x0 = (int)std::lround(cx), y0 = (int)std::lround(cy);
                for (int dy = -r; dy <= r; ++dy)
                    for (int dx = -r; dx <= r; ++dx)
                        if (dx * dx + dy * dy <= r * r) blend(x0 + dx, y0 + dy, a);
            }
            void line(float x0, float y0, float x1, float y1, float a)
            {
                const float len = std::fmax(std::fabs(x1 - x0), std::fabs(y1 - y0));
                const int steps = (int)std::ceil(len);
                if (steps == 0)
                {
                    blend((int)std::lround(x0), (int)std::lround(y0), a);
                    return;
                }
                for (int k = 0; k <= steps; ++k)
                {
                    const float t = (float)k / steps;
                    blend((int)std::lround(x0 + t * (x1 - x0)), (int)std::lround(y0 + t * (y1 - y0)), a);
                }
            }
        };

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

    // PNG of the layout: the layout's bounding box fitted to the image, white background,
    // nodes as discs of radius 3 with opacity 10000/V, edges as lines with opacity
    // 100000/E, both black and blended (src/RPGraphLayout.cpp:170-207; the reference draws
    // through the pngwriter library, which is not part of its checkout, so the pixels
    // themselves are this build's own rasterisation).
    bool GraphLayout::writeToPNG(const int image_w, const int image_h, std::string path)
    {
        if (image_w <= 0 || image_h <= 0) return false;
        const nid_t V = graph.num_nodes();
        const float xr = getXRange(), yr = getYRange();
        const Coordinate c = getCenter();
        const float x0 = c.x - xr / 2.0, y0 = c.y - yr / 2.0;
        const float sx = xr > 0 ? image_w / xr : 1.0f, sy = yr > 0 ? image_h / yr : 1.0f;
        const float node_a = std::fmin(1.0f, (float)(10000.0 / std::max<nid_t>(V, 1)));
        const float edge_a = std::fmin(1.0f, (float)(100000.0 / std::max<nid_t>(graph.num_edges(), 1)));
        Canvas cv(image_w, image_h);
        for (nid_t n1 = 0; n1 < V; ++n1)
        {
            const float ax = (getX(n1) - x0) * sx, ay = (getY(n1) - y0) * sy;
            cv.disc(ax, ay, 3, node_a);
            for (nid_t n2 : graph.neighbors_with_geq_id(n1))
                cv.line(ax, ay, (getX(n2) - x0) * sx, (getY(n2) - y0) * sy, edge_a);
        }
        std::vector<unsigned char> raw;
        raw.reserve((size_t)image_h * (3 * image_w + 1));
        for (int y = 0; y < image_h; ++y)
        {
            raw.push_back(0);
            for (int x = 0; x < image_w; ++x)
            {
                const unsigned char v = (unsigned char)std::lround(255.0f * cv.keep[(size_t)y * image_w + x]);
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
