// Tile choices of the tiled kernels and the table the autotuner fills.
#pragma once

#include <cstdint>

namespace gt355 {

    // Every tile shape instantiated for hdiff: X(TI, TJ, KC).  TI*TJ must be a multiple of 256.
#define GT355_HDIFF_TILES(X)                                                                                     \
    X(64, 4, 4) X(64, 4, 5) X(64, 4, 8) X(64, 4, 10) X(64, 4, 16)                                                  \
    X(64, 8, 4) X(64, 8, 5) X(64, 8, 8) X(64, 8, 10) X(64, 8, 16)                                                  \
    X(64, 16, 5) X(64, 16, 8)                                                                                      \
    X(128, 2, 8) X(128, 2, 10) X(128, 4, 5) X(128, 4, 8)                                                           \
    X(256, 1, 8) X(256, 1, 16)                                                                                     \
    X(32, 8, 8) X(32, 16, 8)

    // Every register-sweep shape instantiated for hdiff: X(JB, P).
#define GT355_HDIFF_SWEEPS(X) X(4, 4) X(8, 4) X(16, 4) X(16, 8) X(32, 4) X(32, 8)

    // Every column-pair sweep shape instantiated for hdiff: X(JB, P).
#define GT355_HDIFF_SWEEPS2(X) X(4, 4) X(8, 4) X(8, 8) X(16, 4) X(16, 8)

    // One hdiff launch configuration: impl 0 = LDS tile (a, b, c) = (TI, TJ, KC); impl 1 = register sweep
    // (a, b) = (JB, P); impl 2 = column-pair register sweep (a, b) = (JB, P).
    struct hdiff_cfg {
        int impl, a, b, c;
    };

    // The configuration for an hdiff launch, in this order: environment overrides (GT355_TUNE_hdiff_*), the entry
    // the autotuner recorded for (storage size, compute size, ni, nj, nk), the measured default for the storage type.
    hdiff_cfg hdiff_cfg_for(int st_size, int ct_size, int ni, int nj, int nk);
    // the measured default configuration for a storage size and domain width (no tuned entry, no knob)
    hdiff_cfg hdiff_default_cfg(int st_size, int ni);
    // whether a configuration is one of the instantiated ones (hdiff.hip)
    bool hdiff_cfg_instantiated(hdiff_cfg cfg);
    void record_hdiff_cfg(int st_size, int ct_size, int ni, int nj, int nk, hdiff_cfg cfg, double us);

    // A batch's launch configuration, as the batch tuner records it for one list of domains (its signature).
    struct batch_cfg {
        int hjb;     // stencil phase: rows per item (8, 16, 32)
        int hw;      // stencil phase: workgroups per CU (2, 4, 8)
        int w;       // vertical phase (or the single-launch forms): the persistent grid's setting (1..4)
        int deep_nk; // vertical phase, FP32 storage: columns of at least this many levels keep 96 in registers
    };
    bool find_batch_cfg(uint64_t sig, int st_size, int ct_size, int impl, batch_cfg *out);
    void record_batch_cfg(uint64_t sig, int st_size, int ct_size, int impl, batch_cfg cfg, double us);

    // A named integer tuning knob: the environment variable GT355_TUNE_<name> if set, otherwise dflt.  The
    // environment is read once per process; tune_env_refresh (gt355_tune_env_refresh) reads it again.
    int tune_param(const char *name, int dflt);
    void tune_env_refresh();

} // namespace gt355
