// Launch-configuration selection and the autotuner's table.
//
// The defaults were measured on MI355X at 256x256x80.  gt355_tune_hdiff (hdiff.hip) times every instantiated
// configuration for one problem shape and records the fastest here, and gt355_batch_tune (batch.hip) the fastest
// launch configuration of a batch, keyed by the signature of its domain list; gt355_tune_save / gt355_tune_load
// persist the table as text, one entry per line:
//   hdiff <storage bytes> <compute bytes> <ni> <nj> <nk> <impl> <a> <b> <c> <microseconds>
//   batch <signature, hex> <storage bytes> <compute bytes> <form> <hjb> <hw> <w> <deep_nk> <microseconds>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "common.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {
        using tune_key = std::tuple<int, int, int, int, int>;
        struct entry {
            hdiff_cfg cfg;
            double us;
        };
        std::mutex g_tune_mutex;
        std::map<tune_key, entry> g_hdiff;
        using batch_key = std::tuple<uint64_t, int, int, int>;
        struct batch_entry {
            batch_cfg cfg;
            double us;
        };
        std::map<batch_key, batch_entry> g_batch;
    } // namespace

    bool find_batch_cfg(uint64_t sig, int st_size, int ct_size, int impl, batch_cfg *out) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_batch.find(batch_key{sig, st_size, ct_size, impl});
        if (it == g_batch.end())
            return false;
        *out = it->second.cfg;
        return true;
    }

    void record_batch_cfg(uint64_t sig, int st_size, int ct_size, int impl, batch_cfg cfg, double us) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_batch[batch_key{sig, st_size, ct_size, impl}] = batch_entry{cfg, us};
    }

    // The GT355_TUNE_* knobs, read from the environment once (at the first lookup) and again only on
    // gt355_tune_env_refresh: a launch looks its knobs up in this snapshot, with no getenv and no allocation.
    namespace {
        struct knob {
            char name[48];
            int value;
        };
        struct knob_table {
            knob k[64];
            int n = 0;
        };
        std::mutex g_knob_mutex;
        knob_table g_knobs;
        bool g_knobs_read = false;

        void read_knobs_locked() {
            g_knobs.n = 0;
            const size_t plen = std::strlen("GT355_TUNE_");
            for (char **e = environ; e && *e; ++e) {
                if (std::strncmp(*e, "GT355_TUNE_", plen) != 0)
                    continue;
                const char *eq = std::strchr(*e, '=');
                if (!eq || g_knobs.n == 64)
                    continue;
                const size_t len = (size_t)(eq - *e) - plen;
                if (len == 0 || len >= sizeof(g_knobs.k[0].name))
                    continue;
                knob &kn = g_knobs.k[g_knobs.n++];
                std::memcpy(kn.name, *e + plen, len);
                kn.name[len] = 0;
                kn.value = std::atoi(eq + 1);
            }
            g_knobs_read = true;
        }
    } // namespace

    int tune_param(const char *name, int dflt) {
        std::lock_guard<std::mutex> lock(g_knob_mutex);
        if (!g_knobs_read)
            read_knobs_locked();
        for (int n = 0; n < g_knobs.n; ++n)
            if (std::strcmp(g_knobs.k[n].name, name) == 0)
                return g_knobs.k[n].value;
        return dflt;
    }

    void tune_env_refresh() {
        std::lock_guard<std::mutex> lock(g_knob_mutex);
        read_knobs_locked();
    }

    hdiff_cfg hdiff_default_cfg(int st_size, int ni) {
        // measured defaults (MI355X, x80 levels, autotuner winners): FP32 storage takes the column-pair sweep with
        // 8-row strips (256x256: 13.5 us, 512x512: 41 us); FP64 storage the one-column sweep with 8-row strips at
        // 256x256 (22.2 us) and the column-pair sweep from 512 columns on (88 us against 97 us)
        return st_size == 4 ? hdiff_cfg{2, 8, 4, 0} : ni >= 512 ? hdiff_cfg{2, 8, 4, 0} : hdiff_cfg{1, 8, 4, 0};
    }

    hdiff_cfg hdiff_cfg_for(int st_size, int ct_size, int ni, int nj, int nk) {
        hdiff_cfg c = hdiff_default_cfg(st_size, ni);
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            auto it = g_hdiff.find(tune_key{st_size, ct_size, ni, nj, nk});
            if (it != g_hdiff.end())
                c = it->second.cfg;
        }
        const int impl = tune_param("hdiff_impl", c.impl);
        if (impl != c.impl)
            c = impl == 2 ? hdiff_cfg{2, 8, 4, 0} : impl == 1 ? hdiff_cfg{1, 16, 4, 0} : hdiff_cfg{0, 64, 8, 5};
        if (c.impl >= 1) {
            c.a = tune_param("hdiff_jb", c.a);
            c.b = tune_param("hdiff_p", c.b);
        } else {
            c.a = tune_param("hdiff_ti", c.a);
            c.b = tune_param("hdiff_tj", c.b);
            c.c = tune_param("hdiff_kc", c.c);
        }
        return c;
    }

    void record_hdiff_cfg(int st_size, int ct_size, int ni, int nj, int nk, hdiff_cfg cfg, double us) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_hdiff[tune_key{st_size, ct_size, ni, nj, nk}] = entry{cfg, us};
    }

} // namespace gt355

using namespace gt355;

extern "C" int gt355_tune_save(const char *path) {
    GT355_CHECK(path != nullptr, "gt355_tune_save: null path");
    FILE *f = std::fopen(path, "w");
    GT355_CHECK(f != nullptr, "gt355_tune_save: cannot open %s", path);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    for (const auto &kv : g_hdiff) {
        const auto &k = kv.first;
        const auto &e = kv.second;
        std::fprintf(f, "hdiff %d %d %d %d %d %d %d %d %d %.3f\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
            std::get<3>(k), std::get<4>(k), e.cfg.impl, e.cfg.a, e.cfg.b, e.cfg.c, e.us);
    }
    for (const auto &kv : g_batch) {
        const auto &k = kv.first;
        const auto &e = kv.second;
        std::fprintf(f, "batch %016llx %d %d %d %d %d %d %d %.3f\n", (unsigned long long)std::get<0>(k),
            std::get<1>(k), std::get<2>(k), std::get<3>(k), e.cfg.hjb, e.cfg.hw, e.cfg.w, e.cfg.deep_nk, e.us);
    }
    std::fclose(f);
    return GT355_OK;
}

extern "C" int gt355_tune_load(const char *path, int *entries) {
    GT355_CHECK(path != nullptr, "gt355_tune_load: null path");
    FILE *f = std::fopen(path, "r");
    GT355_CHECK(f != nullptr, "gt355_tune_load: cannot open %s", path);
    char line[256];
    int st, ct, ni, nj, nk, impl, a, b, c;
    unsigned long long sig;
    double us;
    int n = 0;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    while (std::fgets(line, sizeof line, f)) {
        if (std::sscanf(line, "hdiff %d %d %d %d %d %d %d %d %d %lf", &st, &ct, &ni, &nj, &nk, &impl, &a, &b, &c,
                &us) == 10) {
            g_hdiff[tune_key{st, ct, ni, nj, nk}] = entry{hdiff_cfg{impl, a, b, c}, us};
            ++n;
        } else if (std::sscanf(line, "batch %llx %d %d %d %d %d %d %d %lf", &sig, &st, &ct, &impl, &a, &b, &c, &nk,
                       &us) == 9) {
            g_batch[batch_key{(uint64_t)sig, st, ct, impl}] = batch_entry{batch_cfg{a, b, c, nk}, us};
            ++n;
        }
    }
    std::fclose(f);
    if (entries)
        *entries = n;
    return GT355_OK;
}

// The table by kernel key.  "hdiff": shape {storage bytes, compute bytes, ni, nj, nk}, cfg {impl, a, b, c};
// "batch": shape {signature, storage bytes, compute bytes, form}, cfg {hjb, hw, w, deep_nk}.
namespace {
    int tune_key_of(const char *key, const int64_t *shape, int n, const char *what, int *kind) {
        GT355_CHECK(key != nullptr && shape != nullptr, "%s: null argument", what);
        if (std::strcmp(key, "hdiff") == 0) {
            GT355_CHECK(n == 5, "%s: an hdiff shape has 5 values (storage bytes, compute bytes, ni, nj, nk), not %d",
                what, n);
            *kind = 0;
        } else if (std::strcmp(key, "batch") == 0) {
            GT355_CHECK(n == 4, "%s: a batch shape has 4 values (signature, storage bytes, compute bytes, form), not %d",
                what, n);
            *kind = 1;
        } else {
            GT355_CHECK(false, "%s: unknown kernel key '%s' (hdiff, batch)", what, key);
        }
        return GT355_OK;
    }
} // namespace

extern "C" int gt355_tune_lookup(const char *kernel_key, const int64_t *shape, int n, int32_t cfg[4], double *us,
    int *found) {
    int kind = 0;
    if (int rc = tune_key_of(kernel_key, shape, n, "gt355_tune_lookup", &kind))
        return rc;
    GT355_CHECK(cfg != nullptr && found != nullptr, "gt355_tune_lookup: null output");
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    *found = 0;
    if (kind == 0) {
        auto it = g_hdiff.find(tune_key{(int)shape[0], (int)shape[1], (int)shape[2], (int)shape[3], (int)shape[4]});
        if (it != g_hdiff.end()) {
            const hdiff_cfg &c = it->second.cfg;
            cfg[0] = c.impl, cfg[1] = c.a, cfg[2] = c.b, cfg[3] = c.c;
            if (us)
                *us = it->second.us;
            *found = 1;
        }
    } else {
        auto it = g_batch.find(batch_key{(uint64_t)shape[0], (int)shape[1], (int)shape[2], (int)shape[3]});
        if (it != g_batch.end()) {
            const batch_cfg &c = it->second.cfg;
            cfg[0] = c.hjb, cfg[1] = c.hw, cfg[2] = c.w, cfg[3] = c.deep_nk;
            if (us)
                *us = it->second.us;
            *found = 1;
        }
    }
    return GT355_OK;
}

extern "C" int gt355_tune_store(const char *kernel_key, const int64_t *shape, int n, const int32_t cfg[4], double us) {
    int kind = 0;
    if (int rc = tune_key_of(kernel_key, shape, n, "gt355_tune_store", &kind))
        return rc;
    GT355_CHECK(cfg != nullptr, "gt355_tune_store: null configuration");
    if (kind == 0) {
        const hdiff_cfg c{cfg[0], cfg[1], cfg[2], cfg[3]};
        GT355_CHECK(hdiff_cfg_instantiated(c), "gt355_tune_store: hdiff configuration impl=%d (%d,%d,%d) is not "
                                                "instantiated", c.impl, c.a, c.b, c.c);
        GT355_CHECK((shape[0] == 4 || shape[0] == 8) && (shape[1] == 4 || shape[1] == 8) && shape[2] > 0 &&
                        shape[3] > 0 && shape[4] > 0,
            "gt355_tune_store: bad hdiff shape");
        record_hdiff_cfg((int)shape[0], (int)shape[1], (int)shape[2], (int)shape[3], (int)shape[4], c, us);
    } else {
        GT355_CHECK((cfg[0] == 8 || cfg[0] == 16 || cfg[0] == 32) && cfg[1] >= 1 && cfg[1] <= 8 && cfg[2] >= 1 &&
                        cfg[2] <= 4 && cfg[3] >= 2,
            "gt355_tune_store: bad batch configuration {%d, %d, %d, %d}", cfg[0], cfg[1], cfg[2], cfg[3]);
        record_batch_cfg((uint64_t)shape[0], (int)shape[1], (int)shape[2], (int)shape[3],
            batch_cfg{cfg[0], cfg[1], cfg[2], cfg[3]}, us);
    }
    return GT355_OK;
}

extern "C" int gt355_tune_env_refresh(void) {
    gt355::tune_env_refresh();
    return GT355_OK;
}

extern "C" int gt355_tune_clear(void) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_hdiff.clear();
    g_batch.clear();
    return GT355_OK;
}
