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

    hdiff_cfg hdiff_cfg_for(int st_size, int ct_size, int ni, int nj, int nk) {
        // measured defaults (MI355X, x80 levels, autotuner winners): FP32 storage takes the column-pair sweep with
        // 8-row strips (256x256: 13.5 us, 512x512: 41 us); FP64 storage the one-column sweep with 8-row strips at
        // 256x256 (22.2 us) and the column-pair sweep from 512 columns on (88 us against 97 us)
        hdiff_cfg c = st_size == 4 ? hdiff_cfg{2, 8, 4, 0} : ni >= 512 ? hdiff_cfg{2, 8, 4, 0} : hdiff_cfg{1, 8, 4, 0};
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
