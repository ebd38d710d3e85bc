// Tile choices of the tiled kernels and the table the autotuner fills.
#pragma once

#include <cstdint>

namespace gt355 {

    struct hdiff_tile {
        int ti, tj, kc;
    };

    // Every tile shape instantiated for hdiff: X(TI, TJ, KC).  TI*TJ must be a multiple of 256.
#define GT355_HDIFF_TILES(X)                                                                                     \
    X(64, 4, 4) X(64, 4, 5) X(64, 4, 8) X(64, 4, 10) X(64, 4, 16)                                                  \
    X(64, 8, 4) X(64, 8, 5) X(64, 8, 8) X(64, 8, 10) X(64, 8, 16)                                                  \
    X(64, 16, 5) X(64, 16, 8)                                                                                      \
    X(128, 2, 8) X(128, 2, 10) X(128, 4, 5) X(128, 4, 8)                                                           \
    X(256, 1, 8) X(256, 1, 16)                                                                                     \
    X(32, 8, 8) X(32, 16, 8)

    // Tile for an hdiff launch: the tuned entry for (storage size, compute size, ni, nj, nk) if the autotuner
    // recorded one, otherwise a heuristic default.
    hdiff_tile hdiff_tile_for(int st_size, int ct_size, int ni, int nj, int nk);

    // A named integer tuning knob: the environment variable GT355_TUNE_<name> if set, otherwise dflt.
    int tune_param(const char *name, int dflt);

} // namespace gt355
