// Batched evaluation of many resident metrics (batch.hip): one persistent
// launch per pass over a tile queue that spans every pair, the whole
// evaluation captured once as a HIP graph.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"

namespace sitkhip {

struct Metric;
struct BatchEngine;

// Per-pair result of one batched evaluation (value: -MI or mean squares,
// DBL_MAX for a degenerate pair; moments: 12 Mattes or 14 mean-squares).
struct BatchOutput {
    double value;
    size_t valid;
    double moments[14];
};

struct BatchTiling {
    size_t tile = 0;     // samples per tile (0: default)
    unsigned grid = 0;   // persistent workgroups (0: default)
};

BatchEngine *batch_create(Metric *const *metrics, size_t n);
void batch_destroy(BatchEngine *b);
size_t batch_size(const BatchEngine &b);
void batch_set_tiling(BatchEngine &b, const BatchTiling &t);
BatchTiling batch_tiling(const BatchEngine &b);
// stages: n x 12 doubles (one point-map stage per pair); active: n flags or
// nullptr (all). Inactive pairs cost nothing and keep their previous output.
void batch_evaluate(BatchEngine &b, const double *stages, const uint8_t *active,
                    std::vector<BatchOutput> &out);
// The same evaluation in two halves, so the host can work while the device
// evaluates: batch_launch enqueues it and returns (the stages and flags are
// copied; the engine must not be launched again before batch_collect),
// batch_collect waits for that evaluation only and reads its outputs.
// batch_prepare builds the engine's graph ahead of the first launch.
void batch_prepare(BatchEngine &b);
void batch_launch(BatchEngine &b, const double *stages, const uint8_t *active);
void batch_collect(BatchEngine &b, std::vector<BatchOutput> &out);
// Times every candidate tiling on `stages` and keeps the fastest.
BatchTiling batch_autotune(BatchEngine &b, const double *stages, double *best_ms);

}  // namespace sitkhip
