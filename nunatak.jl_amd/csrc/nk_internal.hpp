// nk_internal.hpp — library-internal helpers shared by the solver and the server (not part of the C ABI).
#pragma once
#include <cstdint>

#include "../../include/nunatak_hip.h"

#define NK_INTERNAL __attribute__((visibility("hidden")))

extern "C" {
// Packed extents (x, y, σ) of a field for a grid and physics, as nk_field_shape reports them for a handle.
NK_INTERNAL void nk_internal_field_extents(const nk_grid *grid, const nk_physics *physics, int id, int32_t shape[3]);
// Zero every field plane whose id is not marked in keep[NK_FIELD_COUNT], in one pass on the handle's stream.
NK_INTERNAL nk_status nk_internal_reset_fields(nk_solver *solver, const uint8_t *keep);
}
