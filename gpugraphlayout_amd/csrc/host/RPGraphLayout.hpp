// Node positions of a graph and the layout writers. Mirrors the reference's GraphLayout
// (src/RPGraphLayout.hpp:33-67, src/RPGraphLayout.cpp): uniform random initial
// positions, extent helpers, and PNG / CSV / BIN output. Writers refuse to overwrite an
// existing file (the reference prints and exits, src/RPGraphLayout.cpp:211-215; here they
// print and return false).
#ifndef RPGraphLayout_hpp
#define RPGraphLayout_hpp
#include <string>
#include <vector>
#include "RPCommon.hpp"
#include "RPGraph.hpp"
#include "gfa2.h"

namespace RPGraph
{
    class GraphLayout
    {
    public:
        GraphLayout(UGraph& graph, float width = 10000, float height = 10000);

        UGraph& graph;

        // x then y per node from get_random over [-w/2, w/2] x [-h/2, h/2]
        // (src/RPGraphLayout.cpp:45-52).
        void randomizePositions();

        float getX(nid_t node_id) const { return xy[2 * node_id]; }
        float getY(nid_t node_id) const { return xy[2 * node_id + 1]; }
        void setX(nid_t node_id, float v) { xy[2 * node_id] = v; }
        void setY(nid_t node_id, float v) { xy[2 * node_id + 1] = v; }
        float getXRange(), getYRange(), getSpan();
        float getDistance(nid_t n1, nid_t n2);
        Real2DVector getDistanceVector(nid_t n1, nid_t n2);   // p(n2) - p(n1)
        Coordinate getCoordinate(nid_t node_id) { return Coordinate(getX(node_id), getY(node_id)); }
        Coordinate getCenter();
        void moveNode(nid_t n, Real2DVector v) { setX(n, getX(n) + v.x); setY(n, getY(n) + v.y); }
        void setCoordinates(nid_t n, Coordinate c) { setX(n, c.x); setY(n, c.y); }

        // Interleaved x,y of all nodes (what the engine consumes and returns).
        float* data() { return xy.data(); }
        const float* data() const { return xy.data(); }

        bool writeToPNG(const int image_w, const int image_h, std::string path);
        bool writeToCSV(std::string path);
        bool writeToBin(std::string path);

    protected:
        float width, height;
        float minX(), minY(), maxX(), maxY();

    private:
        std::vector<float> xy;
    };

    // The picture of writeToPNG, shared by the CPU writer and the device rasteriser
    // (gfa2_render): the frame fitting the bounding box to the image (arithmetic of
    // src/RPGraphLayout.cpp:170-185), the CPU rasteriser producing the same bytes as the
    // device one, and the PNG encoder (8-bit grey written as RGB, zlib-deflated).
    gfa2_raster png_frame(float minx, float maxx, float miny, float maxy, int image_w, int image_h, nid_t nodes,
                          nid_t edges);
    void render_gray(const gfa2_raster& f, const UGraph& graph, const float* xy, std::vector<unsigned char>& gray);
    bool write_png_gray(const std::string& path, int image_w, int image_h, const unsigned char* gray);
}

#endif
