// ForceAtlas2 on an MI355X through the engine's C ABI (include/gfa2.h). Drop-in for the
// reference's CUDAForceAtlas2 (src/RPGPUForceAtlas2.hpp:30-68): same constructor and
// doStep/sync_layout contract (the device layout is authoritative between syncs), plus
// load_layout (host to device) and stats(). Errors are reported as exceptions carrying
// the engine's status instead of exiting the process.
#ifndef RPGPUForceAtlas2_hpp
#define RPGPUForceAtlas2_hpp
#include <stdexcept>
#include <string>
#include "RPForceAtlas2.hpp"
#include "gfa2.h"

namespace RPGraph
{
    class EngineError : public std::runtime_error
    {
    public:
        EngineError(int code, const std::string& what);
        int code;
    };

    class GPUForceAtlas2 : public ForceAtlas2
    {
    public:
        GPUForceAtlas2(GraphLayout& layout, bool use_barneshut, bool strong_gravity, float gravity, float scale,
                       int device = 0, int group_size = 32);
        ~GPUForceAtlas2() override;
        GPUForceAtlas2(const GPUForceAtlas2&) = delete;
        GPUForceAtlas2& operator=(const GPUForceAtlas2&) = delete;

        void doStep() override;           // one iteration, then wait (as the reference)
        void doSteps(int n) override;     // n iterations queued back to back, one wait
        void setScale(float s) override;
        void setGravity(float g) override;
        void sync_layout() override;      // device positions -> layout
        void load_layout();               // layout -> device positions
        void use_graph_capture(bool on);
        // The picture of GraphLayout::writeToPNG, rasterised on the device from the device
        // positions (the same bytes as the CPU writer on the same positions).
        bool writeToPNG(int image_w, int image_h, std::string path);
        gfa2_stats stats();
        gfa2_handle handle() { return h; }

    private:
        gfa2_handle h = nullptr;
        gfa2_params params;
        void check(int rc, const char* what);
        // Pushes the public ForceAtlas2 flags (use_barneshut: exact repulsion when false,
        // use_linlog, prevent_overlap: unsupported) to the engine when a caller changed them.
        void push_flags();
    };

    // The reference's class name, for callers written against it.
    typedef GPUForceAtlas2 CUDAForceAtlas2;
}

#endif
