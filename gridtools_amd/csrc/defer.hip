// The deferred-run scope: gt355_batch_begin / gt355_batch_flush / gt355_batch_end.
//
// The reference's run() launches a spec at once on one grid (stencil/frontend/run.hpp:34-60); its only batching,
// expandable_run (stencil/frontend/expandable_run.hpp:80-183), needs every member on the same grid.  Inside a
// scope, gt355_hdiff / gt355_vadv / gt355_hdiff_vadv calls on the scope's stream are checked by their entry points
// and recorded here instead.  At flush, every recorded part (an hdiff or a vadv) gets a slot in a sequence of
// launches, stencil phase and vertical phase alternating:
//   slot(p) = the first slot of p's phase after the slots of every earlier part p conflicts with,
// where two parts conflict when one writes bytes the other reads or writes (the spans of their access boxes).
// Running the slots in order then gives the results of the calls in recorded order.  The parts of a slot with
// one dtype, precision and dtr run as one two-phase batch (batch.hip) whose domains take that phase only; a part
// alone in its group, or one the batch kernels cannot take (fields of different layouts, a depth outside 2..144
// for the vertical phase), runs on its own entry point (the single-domain kernels, tuned per shape) in its slot.
// The slots run one after another on the scope's stream.  (Running an independent vertical slot concurrently on a
// forked side stream was measured slower: a vertical-phase wavefront holds a SIMD's whole register file, so
// stencil wavefronts placed first delay its single round; hdiff + vadv at 256x256x80 FP64 took 72.6 us per step
// overlapped against 63.3 us in sequence.)  Batches are kept per thread by their exact domain list, so a scope
// that records the same calls again (a timestep loop) only launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace gt355 {
    int batch_create_form(int precision, const gt355_batch_domain *domains, int ndomains, int wg_per_cu,
        gt355_batch **out, int form);

    namespace {
        enum { PART_HDIFF = 0, PART_VADV = 1 }; // also the phase: stencil (even slots), vertical (odd slots)

        // one recorded part; fields: hdiff {out, in, coeff}, vadv {utens_stage, u_stage, wcon, u_pos, utens}
        struct part {
            int kind, precision, ni, nj, nk;
            double dtr;
            gt355_field f[5];
            int slot;
        };
        struct scope {
            void *stream;
            std::vector<part> parts;
        };
        struct cached {
            std::vector<unsigned char> key;
            gt355_batch *batch;
        };
        constexpr size_t CACHE_MAX = 16;
        // a thread's kept batches
        struct thread_state {
            std::list<cached> cache;
        };

        thread_local std::vector<scope> t_scopes;
        thread_local bool t_running = false; // flushing: the entry points launch instead of recording
        thread_local thread_state *t_state = nullptr;
        // every thread's state, so that gt355_finalize can release the batches
        std::mutex g_mu;
        std::vector<thread_state *> g_states;

        thread_state &state() {
            if (!t_state) {
                t_state = new thread_state();
                std::lock_guard<std::mutex> g(g_mu);
                g_states.push_back(t_state);
            }
            return *t_state;
        }

        scope *find_scope(void *stream) {
            for (scope &s : t_scopes)
                if (s.stream == stream)
                    return &s;
            return nullptr;
        }

        // byte span [lo, hi) of a field's access box [ilo, ihi) x [jlo, jhi) x [0, nk)
        struct span {
            uintptr_t lo, hi;
            bool write;
        };
        span span_of(const gt355_field &f, int ilo, int ihi, int jlo, int jhi, int nk, bool write) {
            const int64_t es = f.dtype == GT355_F64 ? 8 : 4;
            const int64_t box[3][2] = {{ilo, ihi - 1}, {jlo, jhi - 1}, {0, nk - 1}};
            int64_t lo = 0, hi = 0;
            for (int d = 0; d < 3; ++d) {
                const int64_t a = box[d][0] * f.stride[d], b = box[d][1] * f.stride[d];
                lo += std::min(a, b);
                hi += std::max(a, b);
            }
            const uintptr_t o = (uintptr_t)f.origin;
            return span{o + (uintptr_t)(lo * es), o + (uintptr_t)((hi + 1) * es), write};
        }
        void spans_of(const part &p, span out[5], int *n) {
            const int ni = p.ni, nj = p.nj, nk = p.nk;
            if (p.kind == PART_HDIFF) {
                out[0] = span_of(p.f[0], 0, ni, 0, nj, nk, true);
                out[1] = span_of(p.f[1], -2, ni + 2, -2, nj + 2, nk, false);
                out[2] = span_of(p.f[2], 0, ni, 0, nj, nk, false);
                *n = 3;
            } else {
                out[0] = span_of(p.f[0], 0, ni, 0, nj, nk, true);
                out[1] = span_of(p.f[1], 0, ni, 0, nj, nk, false);
                out[2] = span_of(p.f[2], 0, ni + 1, 0, nj, nk, false);
                out[3] = span_of(p.f[3], 0, ni, 0, nj, nk, false);
                out[4] = span_of(p.f[4], 0, ni, 0, nj, nk, false);
                *n = 5;
            }
        }
        bool conflict(const part &a, const part &b) {
            span sa[5], sb[5];
            int na, nb;
            spans_of(a, sa, &na);
            spans_of(b, sb, &nb);
            for (int x = 0; x < na; ++x)
                for (int y = 0; y < nb; ++y)
                    if ((sa[x].write || sb[y].write) && sa[x].lo < sb[y].hi && sb[y].lo < sa[x].hi)
                        return true;
            return false;
        }

        bool same_layout(const gt355_field &a, const gt355_field &b) {
            return a.stride[0] == 1 && b.stride[0] == 1 && a.stride[1] == b.stride[1] && a.stride[2] == b.stride[2];
        }
        // whether the batch kernels take the part (else it runs on its own entry point)
        bool batchable(const part &p) {
            const int64_t es = p.f[0].dtype == GT355_F64 ? 8 : 4, lim = int64_t(1) << 32;
            const gt355_field &u = p.f[p.kind == PART_HDIFF ? 0 : 1];
            if (((int64_t)(p.nk - 1) * u.stride[2] + (int64_t)(p.nj + 3) * u.stride[1] + p.ni + 4) * es >= lim)
                return false;
            if (p.kind == PART_HDIFF)
                return same_layout(p.f[0], p.f[1]) && same_layout(p.f[0], p.f[2]) && p.f[0].origin != p.f[1].origin;
            const gt355_field &w = p.f[2];
            return p.nk >= 2 && p.nk <= 144 && same_layout(p.f[0], p.f[1]) && same_layout(p.f[0], p.f[3]) &&
                   same_layout(p.f[0], p.f[4]) && w.stride[0] == 1 &&
                   ((int64_t)(p.nk - 1) * w.stride[2] + (int64_t)(p.nj + 1) * w.stride[1] + p.ni + 1) * es < lim;
        }

        gt355_batch_domain domain_of(const part &p) {
            gt355_batch_domain d;
            std::memset(&d, 0, sizeof d);
            if (p.kind == PART_HDIFF) {
                d.u_stage = p.f[0];
                d.in = p.f[1];
                d.coeff = p.f[2];
                d.phases = 1;
            } else {
                d.utens_stage = p.f[0];
                d.u_stage = p.f[1];
                d.wcon = p.f[2];
                d.u_pos = p.f[3];
                d.utens = p.f[4];
                d.phases = 2;
            }
            d.ni = p.ni, d.nj = p.nj, d.nk = p.nk;
            return d;
        }

        // the batch for a domain list: from the calling thread's cache, else built and cached
        int batch_for(std::vector<gt355_batch_domain> const &doms, int precision, gt355_batch **out) {
            std::list<cached> *t_cache = &state().cache;
            const int dev = current_device();
            std::vector<unsigned char> key(sizeof(int) * 2 + sizeof(gt355_batch_domain) * doms.size());
            std::memcpy(key.data(), &dev, sizeof(int));
            std::memcpy(key.data() + sizeof(int), &precision, sizeof(int));
            std::memcpy(key.data() + 2 * sizeof(int), doms.data(), sizeof(gt355_batch_domain) * doms.size());
            for (auto it = t_cache->begin(); it != t_cache->end(); ++it)
                if (it->key == key) {
                    t_cache->splice(t_cache->begin(), *t_cache, it); // most recently used first
                    *out = t_cache->front().batch;
                    return GT355_OK;
                }
            gt355_batch *b = nullptr;
            int rc = batch_create_form(precision, doms.data(), (int)doms.size(), 2, &b, 3);
            if (rc)
                return rc;
            t_cache->push_front(cached{std::move(key), b});
            if (t_cache->size() > CACHE_MAX) {
                // the evicted batch's launches may still be queued: wait for them before its tables go
                GT355_HIP(hipDeviceSynchronize());
                gt355_batch_destroy(t_cache->back().batch);
                t_cache->pop_back();
            }
            *out = b;
            return GT355_OK;
        }

        int run_direct(const part &p, void *stream) {
            if (p.kind == PART_HDIFF)
                return gt355_hdiff(p.precision, &p.f[0], &p.f[1], &p.f[2], p.ni, p.nj, p.nk, stream);
            return gt355_vadv(
                p.precision, &p.f[0], &p.f[1], &p.f[2], &p.f[3], &p.f[4], p.dtr, p.ni, p.nj, p.nk, stream);
        }

        // the parts of slot s on `stream`: groups of one dtype, precision and dtr (the vertical phase takes dtr as
        // a launch argument) as batches, a group of one part and a part the batch kernels cannot take directly
        int run_slot(std::vector<part> const &parts, int s, void *stream) {
            std::vector<int> done(parts.size(), 0);
            int rc = GT355_OK;
            for (size_t p = 0; p < parts.size() && rc == GT355_OK; ++p) {
                if (parts[p].slot != s || done[p])
                    continue;
                std::vector<size_t> group{p};
                done[p] = 1;
                if (batchable(parts[p]))
                    for (size_t q = p + 1; q < parts.size(); ++q) {
                        const part &a = parts[p], &b = parts[q];
                        if (b.slot != s || done[q] || !batchable(b) || b.precision != a.precision ||
                            b.f[0].dtype != a.f[0].dtype || (a.kind == PART_VADV && b.dtr != a.dtr))
                            continue;
                        done[q] = 1;
                        group.push_back(q);
                    }
                if (group.size() == 1) {
                    rc = run_direct(parts[p], stream);
                    continue;
                }
                std::vector<gt355_batch_domain> doms;
                for (size_t q : group)
                    doms.push_back(domain_of(parts[q]));
                gt355_batch *b = nullptr;
                rc = batch_for(doms, parts[p].precision, &b);
                if (rc == GT355_OK)
                    rc = gt355_batch_run(b, parts[p].kind == PART_VADV ? parts[p].dtr : 0.0, stream);
            }
            return rc;
        }

        // runs and clears what the scope recorded
        int run_scope(scope &sc) {
            std::vector<part> parts;
            parts.swap(sc.parts);
            if (parts.empty())
                return GT355_OK;
            int nslots = 0;
            for (size_t p = 0; p < parts.size(); ++p) {
                int lo = 0;
                for (size_t q = 0; q < p; ++q)
                    if (parts[q].slot + 1 > lo && conflict(parts[p], parts[q]))
                        lo = parts[q].slot + 1;
                parts[p].slot = lo + ((lo & 1) != parts[p].kind ? 1 : 0);
                nslots = std::max(nslots, parts[p].slot + 1);
            }
            t_running = true;
            int rc = GT355_OK;
            for (int s = 0; s < nslots && rc == GT355_OK; ++s)
                rc = run_slot(parts, s, sc.stream);
            t_running = false;
            return rc;
        }
    } // namespace

    bool defer_active(void *stream) { return !t_running && !t_scopes.empty() && find_scope(stream) != nullptr; }

    int defer_hdiff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni,
        int nj, int nk, void *stream) {
        part p{};
        p.kind = PART_HDIFF, p.precision = precision, p.ni = ni, p.nj = nj, p.nk = nk, p.dtr = 0;
        p.f[0] = *out, p.f[1] = *in, p.f[2] = *coeff;
        find_scope(stream)->parts.push_back(p);
        return GT355_OK;
    }
    int defer_vadv(int precision, const gt355_field *utens_stage, const gt355_field *u_stage, const gt355_field *wcon,
        const gt355_field *u_pos, const gt355_field *utens, double dtr, int ni, int nj, int nk, void *stream) {
        part p{};
        p.kind = PART_VADV, p.precision = precision, p.ni = ni, p.nj = nj, p.nk = nk, p.dtr = dtr;
        p.f[0] = *utens_stage, p.f[1] = *u_stage, p.f[2] = *wcon, p.f[3] = *u_pos, p.f[4] = *utens;
        find_scope(stream)->parts.push_back(p);
        return GT355_OK;
    }

    int defer_flush(void *stream) {
        if (t_running || t_scopes.empty())
            return GT355_OK;
        scope *sc = find_scope(stream);
        return sc ? run_scope(*sc) : GT355_OK;
    }
    int defer_flush_all() {
        if (t_running)
            return GT355_OK;
        for (scope &sc : t_scopes)
            if (int rc = run_scope(sc))
                return rc;
        return GT355_OK;
    }

    void release_deferred() {
        std::lock_guard<std::mutex> g(g_mu);
        for (thread_state *t : g_states) {
            for (cached &e : t->cache)
                gt355_batch_destroy(e.batch);
            t->cache.clear();
        }
    }
} // namespace gt355

using namespace gt355;

extern "C" int gt355_batch_begin(void *stream) {
    GT355_CHECK(!t_running, "gt355_batch_begin: called while a scope's calls run");
    GT355_CHECK(find_scope(stream) == nullptr, "gt355_batch_begin: a scope is already open on this stream");
    t_scopes.push_back(scope{stream, {}});
    return GT355_OK;
}

extern "C" int gt355_batch_flush(void *stream) { return defer_flush(stream); }

extern "C" int gt355_batch_end(void *stream) {
    GT355_CHECK(!t_running, "gt355_batch_end: called while a scope's calls run");
    scope *sc = find_scope(stream);
    GT355_CHECK(sc != nullptr, "gt355_batch_end: no scope is open on this stream");
    const int rc = run_scope(*sc);
    t_scopes.erase(t_scopes.begin() + (sc - t_scopes.data()));
    return rc;
}
