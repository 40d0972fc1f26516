// quill_gpu.h -- the binding a QUILL build adds to run its per-step loop on
// one MI355X through libquillk.so.  Compiled inside the reference's tree
// (it includes its include/main.h) next to src/main.cpp, whose globals it
// reads; see quill_gpu.cpp.
#pragma once

#include "main.h"      // spatial_region (include/main.h:17-170)
#include "quillk.h"

struct quill_gpu_options {
    int precision = QK_FP32;          // QK_FP32 or QK_FP64
    int deposit = QK_DEPOSIT_TILED;
    int sort_every = 20;              // cell-tile sort at least every 20 steps ...
    double sort_fallback = 0.01;      // ... and sooner once 1 % left their tile windows
    int device = 0;
};

// After main() has set the region up (init, create_arrays, fields, films,
// f_init_boundaries and interpolate_be: src/main.cpp:2275-2289): a context
// holding the region's fields and every particle of its cell lists,
// configured from main()'s globals.  nullptr on failure (qk_last_error()).
qk_ctx *quill_gpu_create(spatial_region &sr, const quill_gpu_options &opt);

// The loop body of src/main.cpp:2319-2332 (birth_from_vacuum, padvance,
// compute_N, compute_energy, fadvance) for one step on the GPU; the step's
// diagnostics land in the region's public N_e .. energy_ph, N_qp_*, ienergy.
int quill_gpu_step(qk_ctx *ctx, spatial_region &sr);

// Fields back into the region's ce, cb, cj, cbe (before the reference's
// writers run at an output step).
int quill_gpu_fields_to_region(qk_ctx *ctx, spatial_region &sr);
