// Base of layout algorithms (src/RPLayoutAlgorithm.hpp:31-39), with a virtual destructor
// (the reference's is not, so deleting through the base is undefined there).
#ifndef RPLayoutAlgorithm_hpp
#define RPLayoutAlgorithm_hpp
#include "RPGraphLayout.hpp"

namespace RPGraph
{
    class LayoutAlgorithm
    {
    public:
        explicit LayoutAlgorithm(GraphLayout& layout) : layout(layout) {}
        virtual ~LayoutAlgorithm() = default;
        GraphLayout& layout;
        virtual void sync_layout() = 0;   // write the current layout to `layout'
    };
}

#endif
