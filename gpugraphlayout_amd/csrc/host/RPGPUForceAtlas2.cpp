#include "RPGPUForceAtlas2.hpp"

#include <vector>

namespace RPGraph
{
    EngineError::EngineError(int code, const std::string& what)
        : std::runtime_error(what + ": " + gfa2_strerror(code)), code(code)
    {
    }

    void GPUForceAtlas2::check(int rc, const char* what)
    {
        if (rc != GFA2_OK) throw EngineError(rc, what);
    }

    // The constructor stages what src/RPGPUForceAtlas2.cu:51-84 stages (masses deg+1,
    // each undirected edge once) and hands it to the engine, which owns all device state.
    GPUForceAtlas2::GPUForceAtlas2(GraphLayout& layout, bool use_barneshut, bool strong_gravity, float gravity,
                                   float scale, int device, int group_size)
        : ForceAtlas2(layout, use_barneshut, strong_gravity, gravity, scale)
    {
        gfa2_default_params(&params);
        params.use_barneshut = use_barneshut ? 1 : 0;
        params.strong_gravity = strong_gravity ? 1 : 0;
        params.gravity = k_g;
        params.scale = k_r;
        params.theta = theta;
        params.epssq = epssq;
        params.group_size = group_size;
        params.linlog = use_linlog ? 1 : 0;
        params.prevent_overlap = prevent_overlap ? 1 : 0;
        check(gfa2_create(&params, device, &h), "gfa2_create");
        std::vector<int32_t> src, dst;
        layout.graph.coo(src, dst);
        std::vector<float> m(layout.graph.num_nodes());
        for (nid_t n = 0; n < layout.graph.num_nodes(); ++n) m[n] = mass(n);
        int rc = gfa2_set_graph_coo(h, (int32_t)layout.graph.num_nodes(), (int64_t)src.size(), src.data(), dst.data(),
                                    m.data());
        if (rc == GFA2_OK) rc = gfa2_set_positions(h, layout.data());
        if (rc != GFA2_OK)
        {
            gfa2_destroy(h);
            h = nullptr;
            throw EngineError(rc, "GPUForceAtlas2 setup");
        }
    }

    GPUForceAtlas2::~GPUForceAtlas2()
    {
        if (h) gfa2_destroy(h);
    }

    void GPUForceAtlas2::push_flags()
    {
        const int32_t bh = use_barneshut ? 1 : 0, ll = use_linlog ? 1 : 0, po = prevent_overlap ? 1 : 0;
        if (bh == params.use_barneshut && ll == params.linlog && po == params.prevent_overlap) return;
        gfa2_params next = params;
        next.use_barneshut = bh;
        next.linlog = ll;
        next.prevent_overlap = po;
        check(gfa2_set_params(h, &next), "gfa2_set_params");
        params = next;
    }

    void GPUForceAtlas2::doStep()
    {
        push_flags();
        check(gfa2_step(h, 1), "gfa2_step");
        check(gfa2_sync(h), "gfa2_sync");
        iteration++;
    }

    void GPUForceAtlas2::doSteps(int n)
    {
        if (n <= 0) return;
        push_flags();
        check(gfa2_step(h, n), "gfa2_step");
        check(gfa2_sync(h), "gfa2_sync");
        iteration += n;
    }

    void GPUForceAtlas2::setScale(float s)
    {
        ForceAtlas2::setScale(s);
        params.scale = s;
        if (h) check(gfa2_set_params(h, &params), "gfa2_set_params");
    }

    void GPUForceAtlas2::setGravity(float g)
    {
        ForceAtlas2::setGravity(g);
        params.gravity = g;
        if (h) check(gfa2_set_params(h, &params), "gfa2_set_params");
    }

    void GPUForceAtlas2::sync_layout() { check(gfa2_get_positions(h, layout.data()), "gfa2_get_positions"); }

    void GPUForceAtlas2::load_layout() { check(gfa2_set_positions(h, layout.data()), "gfa2_set_positions"); }

    bool GPUForceAtlas2::writeToPNG(int image_w, int image_h, std::string path)
    {
        if (image_w <= 0 || image_h <= 0) return false;
        float bb[4];
        check(gfa2_get_bbox(h, bb), "gfa2_get_bbox");
        const gfa2_raster f = png_frame(bb[0], bb[1], bb[2], bb[3], image_w, image_h, layout.graph.num_nodes(),
                                        layout.graph.num_edges());
        std::vector<unsigned char> gray((size_t)image_w * image_h);
        check(gfa2_render(h, &f, gray.data()), "gfa2_render");
        return write_png_gray(path, image_w, image_h, gray.data());
    }

    void GPUForceAtlas2::use_graph_capture(bool on) { check(gfa2_set_graph_capture(h, on ? 1 : 0), "capture"); }

    gfa2_stats GPUForceAtlas2::stats()
    {
        gfa2_stats s;
        check(gfa2_get_stats(h, &s), "gfa2_get_stats");
        return s;
    }
}
