// HIP graphs of fixed evaluation sequences.
#pragma once

#include <functional>

#include "bfv.h"

namespace fhehip {

// A captured sequence of launches. It owns the stream it was captured on,
// whose scratch buffers the graph's kernels use; replays of one graph must
// therefore not overlap each other.
struct Graph {
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf tmp;  // intermediate buffers of the sequence
    ~Graph();
    void launch(hipStream_t s);
};

// Runs `issue` once on a new stream (which sizes every scratch buffer), then
// captures a second run into a graph.
std::unique_ptr<Graph> capture_graph(std::unique_ptr<Graph> g, const std::function<void(hipStream_t)>& issue);

}  // namespace fhehip
