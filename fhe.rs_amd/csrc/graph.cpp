// HIP graphs of fixed evaluation sequences.
#include "graph.h"

#include "profiler.h"

namespace fhehip {

Graph::~Graph() {
    // teardown: failures here have no caller to report to
    (void)hipDeviceSynchronize();  // replays may run on any stream
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) {
        release_scratch(cap);
        (void)hipStreamDestroy(cap);
    }
}

void Graph::launch(hipStream_t s) { HIPCHK(hipGraphLaunch(exec, s)); }

std::unique_ptr<Graph> capture_graph(std::unique_ptr<Graph> g, const std::function<void(hipStream_t)>& issue) {
    HIPCHK(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
    issue(g->cap);
    HIPCHK(hipStreamSynchronize(g->cap));
    CaptureGuard guard;
    // Thread-local capture: the library's own threads (the request server, the
    // job runner) keep allocating and synchronizing while a graph is captured
    // here; the eager run above sized every scratch buffer the capture uses.
    HIPCHK(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
    try {
        issue(g->cap);
    } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(g->cap, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
    }
    HIPCHK(hipStreamEndCapture(g->cap, &g->graph));
    HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    return g;
}

}  // namespace fhehip
