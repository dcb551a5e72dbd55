// ForceAtlas2 interface (src/RPForceAtlas2.hpp:32-60, src/RPForceAtlas2.cpp:28-77): the
// reference's constants and the doStep/doSteps/setScale/setGravity/mass surface.
// doSteps, setScale and setGravity are virtual here so the GPU class can batch
// iterations without a host synchronisation each and push parameter changes.
#ifndef RPForceAtlas2_hpp
#define RPForceAtlas2_hpp
#include "RPLayoutAlgorithm.hpp"

namespace RPGraph
{
    class ForceAtlas2 : public LayoutAlgorithm
    {
    public:
        // As the reference, construction randomises the layout (src/RPForceAtlas2.cpp:54).
        ForceAtlas2(GraphLayout& layout, bool use_barneshut, bool strong_gravity, float gravity, float scale);
        ~ForceAtlas2() override = default;

        virtual void doStep() = 0;
        virtual void doSteps(int n);
        virtual void setScale(float s);
        virtual void setGravity(float g);
        float mass(nid_t n);   // degree + 1
        bool prevent_overlap, use_barneshut, use_linlog, strong_gravity;

    protected:
        int iteration;
        float k_r, k_g;
        float delta;
        float global_speed;
        float speed_efficiency, jitter_tolerance;
        float k_s, k_s_max;
        float theta, epssq, itolsq;
    };
}

#endif
