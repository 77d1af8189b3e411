// ihg_graph.hip -- HIP graphs of the context's stream: a fixed-step RK4 step
// of an ensemble (or any other sequence of library calls that neither
// synchronises nor allocates) is captured once and replayed as one launch.
// No counterpart in the reference, which launches nothing.
//
// Captured work must not contain memset nodes (the library's RHS and stage
// paths record none) and the live profiler must be off while capturing.
#include <vector>

#include "ihg_internal.h"

struct ihg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

using namespace ihg;

extern "C" int ihg_graph_begin(ihg_ctx *c)
{
    if (!c) return set_error("NULL context");
    if (c->prof_on)
        return set_error("ihg_graph_begin: switch the live profiler off first");
    return hip_check(hipStreamBeginCapture(c->stream,
                                           hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
}

static int instantiate(ihg_graph *g)
{
    return hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr,
                                         0),
                     "hipGraphInstantiate");
}

extern "C" int ihg_graph_destroy(ihg_graph *g)
{
    if (!g) return 0;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return 0;
}

extern "C" int ihg_graph_end(ihg_ctx *c, ihg_graph **out)
{
    if (!c || !out) return set_error("NULL argument");
    *out = nullptr;
    ihg_graph *g = new ihg_graph();
    int rc = hip_check(hipStreamEndCapture(c->stream, &g->graph),
                       "hipStreamEndCapture");
    if (!rc) rc = instantiate(g);
    if (rc) {
        ihg_graph_destroy(g);
        return rc;
    }
    *out = g;
    return 0;
}

extern "C" int ihg_graph_join(ihg_graph *const *parts, int n, ihg_graph **out)
{
    if (!parts || !out || n < 1) return set_error("ihg_graph_join: no parts");
    *out = nullptr;
    ihg_graph *g = new ihg_graph();
    int rc = hip_check(hipGraphCreate(&g->graph, 0), "hipGraphCreate");
    for (int i = 0; !rc && i < n; ++i) {
        if (!parts[i] || !parts[i]->graph) {
            rc = set_error("ihg_graph_join: part %d is empty", i);
            break;
        }
        hipGraphNode_t node;
        // no dependencies: the parts are independent branches
        rc = hip_check(hipGraphAddChildGraphNode(&node, g->graph, nullptr, 0,
                                                 parts[i]->graph),
                       "hipGraphAddChildGraphNode");
    }
    if (!rc) rc = instantiate(g);
    if (rc) {
        ihg_graph_destroy(g);
        return rc;
    }
    *out = g;
    return 0;
}

extern "C" int ihg_graph_launch(ihg_graph *g, ihg_ctx *c)
{
    if (!g || !g->exec || !c) return set_error("NULL argument");
    return hip_check(hipGraphLaunch(g->exec, c->stream), "hipGraphLaunch");
}
