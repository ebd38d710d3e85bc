// Tile selection for the tiled kernels.  Defaults come from measurements on MI355X; gt355_tune_* (the
// autotuner) can override them per problem shape.
#include "tuning.hpp"

#include <cstdlib>
#include <string>

namespace gt355 {

    hdiff_tile hdiff_tile_for(int st_size, int ct_size, int ni, int nj, int nk) {
        (void)st_size;
        (void)ct_size;
        (void)ni;
        (void)nj;
        (void)nk;
        // 64x8 tiles with 5-level chunks: 2048 workgroups at 256x256x80, one full round at 8 waves per SIMD
        hdiff_tile t{64, 8, 5};
        t.ti = tune_param("hdiff_ti", t.ti);
        t.tj = tune_param("hdiff_tj", t.tj);
        t.kc = tune_param("hdiff_kc", t.kc);
        return t;
    }

    int tune_param(const char *name, int dflt) {
        std::string key = std::string("GT355_TUNE_") + name;
        const char *v = std::getenv(key.c_str());
        return v ? std::atoi(v) : dflt;
    }

} // namespace gt355
