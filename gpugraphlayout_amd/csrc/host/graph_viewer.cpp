// graph_viewer: the reference's command line (src/graph_viewer.cpp:46-176) on the MI355X
// engine. Same positional arguments, same srandom(1234) initial layout, same snapshot
// naming (<edgelist basename>_<iteration>.<fmt>) and progress lines. `gpu` and `cuda`
// select the HIP engine; the serial CPU backend is not part of this engine (it is the
// measured CPU baseline). Iterations between two outputs are queued back to back instead
// of synchronising after each.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "RPCommon.hpp"
#include "RPGPUForceAtlas2.hpp"
#include "RPGraph.hpp"
#include "RPGraphLayout.hpp"

static void usage()
{
    fprintf(stderr,
            "Usage: graph_viewer gpu|cpu max_iterations num_snaps sg|wg scale gravity exact|approximate "
            "edgelist_path out_path [png image_w image_h|csv|bin]\n");
}

int main(int argc, const char** argv)
{
    srandom(1234);   // reproducible initial layout (src/graph_viewer.cpp:49)

    if (argc < 10 || (argc > 10 && std::string(argv[10]) == "png" && argc < 13))
    {
        usage();
        return EXIT_FAILURE;
    }
    const std::string backend = argv[1];
    const bool gpu_requested = backend == "gpu" || backend == "cuda";
    int max_iterations, num_screenshots;
    float scale, gravity;
    try
    {
        max_iterations = std::stoi(argv[2]);
        num_screenshots = std::stoi(argv[3]);
        scale = std::stof(argv[5]);
        gravity = std::stof(argv[6]);
    }
    catch (const std::exception&)
    {
        usage();
        return EXIT_FAILURE;
    }
    const bool strong_gravity = std::string(argv[4]) == "sg";
    const bool approximate = std::string(argv[7]) == "approximate";
    const std::string edgelist_path = argv[8];
    const std::string out_path = argv[9];
    std::string out_format = "png";
    int image_w = 1250, image_h = 1250;
    for (int a = 10; a < argc; a++)
    {
        const std::string arg = argv[a];
        if (arg == "png" && a + 2 < argc)
        {
            out_format = "png";
            image_w = std::stoi(argv[a + 1]);
            image_h = std::stoi(argv[a + 2]);
            a += 2;
        }
        else if (arg == "csv") out_format = "csv";
        else if (arg == "bin") out_format = "bin";
    }

    if (!gpu_requested)
    {
        fprintf(stderr, "error: only the GPU engine is built here (use `gpu`).\n");
        return EXIT_FAILURE;
    }
    if (!is_file_exists(edgelist_path))
    {
        fprintf(stderr, "error: No edgelist at %s\n", edgelist_path.c_str());
        return EXIT_FAILURE;
    }
    if (!is_file_exists(out_path))
    {
        fprintf(stderr, "error: No output folder at %s\n", out_path.c_str());
        return EXIT_FAILURE;
    }

    printf("Loading edgelist at '%s'...", edgelist_path.c_str());
    fflush(stdout);
    RPGraph::UGraph graph(edgelist_path);
    if (!graph.ok())
    {
        fprintf(stderr, "\nerror: could not read %s\n", edgelist_path.c_str());
        return EXIT_FAILURE;
    }
    printf("done.\n");
    printf("    fetched %u nodes and %u edges.\n", graph.num_nodes(), graph.num_edges());

    RPGraph::GraphLayout layout(graph);
    try
    {
        RPGraph::GPUForceAtlas2 fa2(layout, approximate, strong_gravity, gravity, scale);
        printf("Started Layout algorithm...\n");
        const int snap_period = num_screenshots > 0 ? (int)std::ceil((float)max_iterations / num_screenshots) : 0;
        const int print_period = std::max(1, (int)std::ceil((float)max_iterations * 0.05));
        int done = 0;
        for (int iteration = 1; iteration <= max_iterations; ++iteration)
        {
            const bool snap = num_screenshots > 0 && (iteration % snap_period == 0 || iteration == max_iterations);
            const bool print = !snap && iteration % print_period == 0;
            if (!snap && !print && iteration != max_iterations) continue;
            fa2.doSteps(iteration - done);
            done = iteration;
            if (snap)
            {
                const std::string name = basename(edgelist_path) + "_" + std::to_string(iteration) + "." + out_format;
                const std::string file = out_path + "/" + name;
                printf("Starting iteration %d (%.2f%%), writing %s...", iteration,
                       100 * (float)iteration / max_iterations, out_format.c_str());
                fflush(stdout);
                fa2.sync_layout();
                bool ok = true;
                if (out_format == "png") ok = fa2.writeToPNG(image_w, image_h, file);   // rasterised on the GPU
                else if (out_format == "csv") ok = layout.writeToCSV(file);
                else ok = layout.writeToBin(file);
                if (!ok) return EXIT_FAILURE;
                printf("done.\n");
            }
            else if (print)
            {
                printf("Starting iteration %d (%.2f%%).\n", iteration, 100 * (float)iteration / max_iterations);
            }
        }
    }
    catch (const RPGraph::EngineError& e)
    {
        fprintf(stderr, "error: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
