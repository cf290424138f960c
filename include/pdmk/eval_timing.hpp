/// Per-evaluation time breakdown, same fields and grouping as
/// cpp/include/pdmk/eval_timing.hpp:10-19:
///   nufft_ms  -- far field;
///   dmk_ms    -- charge permutation + upward pass + plane-wave downward pass;
///   direct_ms -- near field + self correction + de-sort;
///   build_ms  -- 0 for 3-D (the tree is reused).
/// On the GPU the three parts are measured with HIP events on the tree's stream.
#pragma once

namespace pdmk {

struct EvalTiming {
    double nufft_ms = 0.0;
    double dmk_ms = 0.0;
    double direct_ms = 0.0;
    double build_ms = 0.0;
};

}  // namespace pdmk
