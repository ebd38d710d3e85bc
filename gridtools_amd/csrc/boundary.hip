// Boundary conditions and local halo handling for gfx950.
//
// The reference applies a boundary functor to the 26 halo regions of a 3-D box in one kernel whose grid is
// sized by the largest region, so most blocks idle on thin faces (boundaries/apply_gpu.hpp:98-255).  Its
// periodic halos come from GCL: 12 pack kernels per field into 27 message buffers, an MPI self-exchange on a
// periodic process grid, and 12 unpack kernels (gcl/high_level/m_pack*.hpp, m_unpack*.hpp,
// descriptors_manual_gpu.hpp:342-512).
//
// Here every call is one launch over exactly the halo points it touches.  The host lists the selected regions with
// their boxes; each region gets its own run of workgroups, and a workgroup finds its region once (a uniform scan of
// at most 26 block offsets).  Inside a region points are numbered with i fastest in 32-bit arithmetic, and a
// thread's (i, j, k) comes from two divisions by the region's extents done as multiply-high with host-computed
// magic numbers, so no point pays a 64-bit division or a region search.  All fields of a call share the launch.
//
// The i faces of a halo fill write 3 values per row: alone they are partial 32-byte sectors, which the memory turns
// into read-modify-writes.  When the fields are laid out by gt355_storage_layout (so the bytes between a row's right
// halo and the next row's left halo are the layout's own row padding, zero since allocation and outside every
// field), the row's right halo, that padding and the next row's left halo form one aligned run of whole sectors
// (for ni * element size a multiple of 32 bytes), and extra workgroups of the same launch write each such run with
// 16-byte stores: the halo values and zeros.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {

        constexpr int MAX_FIELDS = 4;

        // x / d for 32-bit x as ((mulhi(x, m) + x) >> l): the round-up method of Granlund and Montgomery
        // ("Division by invariant integers using multiplication", PLDI 1994) with l = ceil(log2 d) and
        // m = floor(2^32 (2^l - d) / d) + 1, exact for every 32-bit x.
        struct fastdiv {
            uint32_t d, m, l;
        };
        inline fastdiv make_fastdiv(uint32_t d) {
            uint32_t l = 0;
            while ((uint64_t(1) << l) < d)
                ++l;
            const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1;
            return fastdiv{d, (uint32_t)m, l};
        }
        __device__ __forceinline__ uint32_t div_by(uint32_t x, fastdiv f) {
            return (uint32_t)(((uint64_t)__umulhi(x, f.m) + x) >> f.l);
        }

        constexpr int BC_BLOCK = 256, BC_PER_THREAD = 4, BC_CHUNK = BC_BLOCK * BC_PER_THREAD;

        struct region {
            int32_t lo[3];   // first index of the region (storage coordinates)
            int32_t n[3];    // extent
            int32_t dir[3];  // direction components (-1, 0, 1)
            int64_t start;   // prefix offset in points
            int64_t src;     // unpack: start of the buffer segment this region reads (start, or the segment of the
                             // opposite direction for a local periodic exchange)
            int32_t block0;  // first workgroup of the region
            fastdiv d0, d1;  // division by n[0] and by n[0] * n[1]
        };

        struct region_table {
            region r[26];
            int count;
            int64_t total;
            int blocks; // workgroups of the launch: every region's chunks of BC_CHUNK points
        };

        // the region of this workgroup (uniform) and the point's coordinates
        __device__ __forceinline__ int block_region(const region_table &t) {
            int q = 0;
#pragma unroll 1
            for (int c = 1; c < t.count; ++c)
                if ((int)blockIdx.x >= t.r[c].block0)
                    q = c;
            return q;
        }
        __device__ __forceinline__ void point_of(const region &r, uint32_t loc, int &i, int &j, int &k) {
            const uint32_t row = div_by(loc, r.d0);
            const uint32_t plane = div_by(loc, r.d1);
            i = r.lo[0] + (int)(loc - row * (uint32_t)r.n[0]);
            j = r.lo[1] + (int)(row - plane * (uint32_t)r.n[1]);
            k = r.lo[2] + (int)plane;
        }

        struct bc_args {
            void *dst[MAX_FIELDS];
            const void *src[MAX_FIELDS];
            int nfields;
            int64_t s[3];
            int32_t begin[3], end[3];
            int32_t mode[3];
            double value;
        };

        // The row-gap runs of a halo fill: rows [row0, row0 + nrows) of levels [k0, k0 + nk); run p of a level
        // starts at column end_i + 1 of row row0 + p and holds the right halo (h1 values) of that row, the row
        // padding and the left halo (h0 values) of the next row: len = stride_j - ni elements, in chunks of 16
        // bytes.  Its workgroups follow the region table's.
        struct gap_args {
            int32_t row0, nrows, k0, h0, h1, len, cpr; // cpr: 16-byte chunks per run
            uint32_t total;                            // chunks of every run (0: no gap workgroups)
            fastdiv dcpr, druns;                       // division by cpr and by nrows - 1
        };

        template <class T>
        __device__ __forceinline__ void gap_fill(const gap_args &g, const bc_args &a, int blk) {
            constexpr int V = 16 / sizeof(T);
            typedef T vec __attribute__((ext_vector_type(V)));
            const uint32_t cid = (uint32_t)blk * BC_BLOCK + threadIdx.x;
            if (cid >= g.total)
                return;
            const uint32_t run = div_by(cid, g.dcpr), c = cid - run * (uint32_t)g.cpr;
            const uint32_t lev = div_by(run, g.druns), p = run - lev * (uint32_t)(g.nrows - 1);
            const int r = g.row0 + (int)p, k = g.k0 + (int)lev;
            const int len_i = a.end[0] - a.begin[0] + 1, len_j = a.end[1] - a.begin[1] + 1;
            // the source row of a row: itself inside, clamped (mode 1) or wrapped (mode 2) in the j halo
            auto srow = [&](int j) {
                if (j < a.begin[1])
                    return a.mode[1] == 1 ? a.begin[1] : j + len_j;
                if (j > a.end[1])
                    return a.mode[1] == 1 ? a.end[1] : j - len_j;
                return j;
            };
            const int64_t base = (int64_t)(a.end[0] + 1) + (int64_t)r * a.s[1] + (int64_t)k * a.s[2] + c * V;
            const int64_t right_src = (int64_t)srow(r) * a.s[1] + (int64_t)k * a.s[2];
            const int64_t left_src = (int64_t)srow(r + 1) * a.s[1] + (int64_t)k * a.s[2];
            for (int f = 0; f < a.nfields; ++f) {
                T *fld = (T *)a.dst[f];
                vec v;
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const int el = (int)c * V + e;
                    T x = T(0);
                    if (el < g.h1) {
                        const int col = a.end[0] + 1 + el;
                        x = fld[right_src + (a.mode[0] == 1 ? a.end[0] : col - len_i)];
                    } else if (el >= g.len - g.h0) {
                        const int col = a.begin[0] - (g.len - el);
                        x = fld[left_src + (a.mode[0] == 1 ? a.begin[0] : col + len_i)];
                    }
                    v[e] = x;
                }
                *(vec *)(fld + base) = v;
            }
        }

        // kind 0 zero, 1 value, 2 copy, 3 halo fill (mode per dimension)
        template <class T, int KIND>
        __global__ void __launch_bounds__(BC_BLOCK) bc_kernel(region_table t, bc_args a, gap_args g) {
            if (KIND == 3 && (int)blockIdx.x >= t.blocks) {
                gap_fill<T>(g, a, (int)blockIdx.x - t.blocks);
                return;
            }
            const region &r = t.r[block_region(t)];
            const uint32_t vol = (uint32_t)r.n[0] * (uint32_t)r.n[1] * (uint32_t)r.n[2];
            const uint32_t base = (uint32_t)((int)blockIdx.x - r.block0) * BC_CHUNK + threadIdx.x;
            // the source offset of a halo fill: clamped (mode 1) or wrapped (mode 2) along each nonzero component
            int64_t shift = 0;
            int clamp[3] = {-1, -1, -1};
            if (KIND == 3) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (r.dir[d] == 0)
                        continue;
                    const int len = a.end[d] - a.begin[d] + 1;
                    if (a.mode[d] == 1)
                        clamp[d] = r.dir[d] < 0 ? a.begin[d] : a.end[d];
                    else
                        shift += (int64_t)(r.dir[d] < 0 ? len : -len) * a.s[d];
                }
            }
#pragma unroll
            for (int u = 0; u < BC_PER_THREAD; ++u) {
                const uint32_t loc = base + u * BC_BLOCK;
                if (loc >= vol)
                    break;
                int i, j, k;
                point_of(r, loc, i, j, k);
                const int64_t p = i * a.s[0] + j * a.s[1] + k * a.s[2];
                if (KIND == 3) {
                    const int64_t sp = (clamp[0] >= 0 ? clamp[0] : i) * a.s[0] + (clamp[1] >= 0 ? clamp[1] : j) * a.s[1] +
                                       (clamp[2] >= 0 ? clamp[2] : k) * a.s[2] + shift;
                    for (int f = 0; f < a.nfields; ++f) {
                        T *dst = (T *)a.dst[f];
                        dst[p] = dst[sp];
                    }
                } else {
                    for (int f = 0; f < a.nfields; ++f) {
                        T *dst = (T *)a.dst[f];
                        if (KIND == 0)
                            dst[p] = T(0);
                        else if (KIND == 1)
                            dst[p] = T(a.value);
                        else
                            dst[p] = ((const T *)a.src[f])[p];
                    }
                }
            }
        }

        // pack (UNPACK = false): buffer <- field slab; unpack: field slab <- buffer.  Buffer order: region, field,
        // then the slab with i fastest.
        template <class T, bool UNPACK>
        __global__ void __launch_bounds__(BC_BLOCK) pack_kernel(region_table t, bc_args a, T *buf) {
            const region &r = t.r[block_region(t)];
            const uint32_t vol = (uint32_t)r.n[0] * (uint32_t)r.n[1] * (uint32_t)r.n[2];
            const uint32_t base = (uint32_t)((int)blockIdx.x - r.block0) * BC_CHUNK + threadIdx.x;
#pragma unroll
            for (int u = 0; u < BC_PER_THREAD; ++u) {
                const uint32_t loc = base + u * BC_BLOCK;
                if (loc >= vol)
                    break;
                int i, j, k;
                point_of(r, loc, i, j, k);
                const int64_t p = i * a.s[0] + j * a.s[1] + k * a.s[2];
                for (int f = 0; f < a.nfields; ++f) {
                    T *fld = (T *)a.dst[f];
                    if (UNPACK)
                        fld[p] = buf[r.src * a.nfields + (int64_t)f * vol + loc];
                    else
                        buf[r.start * a.nfields + (int64_t)f * vol + loc] = fld[p];
                }
            }
        }

        // Bounds of direction component d along one dimension, halo_descriptor h = {minus, plus, begin, end,
        // total} (common/halo_descriptor.hpp:72-120).  which: 0 = the halo region (what a BC or unpack writes),
        // 1 = the interior slab a neighbour in that direction needs (what pack reads).
        void dim_range(const int32_t h[5], int d, int which, int32_t &lo, int32_t &n) {
            if (d == 0) {
                lo = h[2];
                n = h[3] - h[2] + 1;
            } else if (which == 0) {
                lo = d < 0 ? h[2] - h[0] : h[3] + 1;
                n = d < 0 ? h[0] : h[1];
            } else {
                // the neighbour below (d = -1) fills its plus halo from our first `plus` interior points
                lo = d < 0 ? h[2] : h[3] - h[0] + 1;
                n = d < 0 ? h[1] : h[0];
            }
        }

        int check_halo(const int32_t halo[3][5], const char *who) {
            for (int d = 0; d < 3; ++d) {
                const int32_t *h = halo[d];
                GT355_CHECK(h[0] >= 0 && h[1] >= 0 && h[2] <= h[3] && h[2] >= h[0] && h[3] + h[1] < h[4],
                    "%s: invalid halo descriptor in dim %d {%d,%d,%d,%d,%d}", who, d, h[0], h[1], h[2], h[3], h[4]);
            }
            return GT355_OK;
        }

        // which: 0 halo regions, 1 pack slabs.  mode (halo fill only): a direction is kept only if every nonzero
        // component has mode != 0.
        region_table make_table(const int32_t halo[3][5], uint32_t dir_mask, int which, const int32_t *mode) {
            region_table t{};
            int64_t start = 0;
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj)
                    for (int dk = -1; dk <= 1; ++dk) {
                        if (di == 0 && dj == 0 && dk == 0)
                            continue;
                        int bit = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
                        if (!((dir_mask >> bit) & 1u))
                            continue;
                        int dd[3] = {di, dj, dk};
                        if (mode) {
                            bool ok = true;
                            for (int d = 0; d < 3; ++d)
                                if (dd[d] != 0 && mode[d] == 0)
                                    ok = false;
                            if (!ok)
                                continue;
                        }
                        region r{};
                        int64_t vol = 1;
                        for (int d = 0; d < 3; ++d) {
                            dim_range(halo[d], dd[d], which, r.lo[d], r.n[d]);
                            r.dir[d] = dd[d];
                            vol *= r.n[d];
                        }
                        if (vol == 0)
                            continue;
                        r.start = start;
                        r.src = start;
                        r.block0 = t.blocks;
                        r.d0 = make_fastdiv((uint32_t)r.n[0]);
                        r.d1 = make_fastdiv((uint32_t)r.n[0] * (uint32_t)r.n[1]);
                        t.blocks += (int)((vol + BC_CHUNK - 1) / BC_CHUNK);
                        start += vol;
                        t.r[t.count++] = r;
                    }
            t.total = start;
            return t;
        }

        // every region's points fit 32-bit numbering, and the workgroups fit one launch
        int check_table(const region_table &t, const char *who) {
            for (int q = 0; q < t.count; ++q)
                GT355_CHECK((int64_t)t.r[q].n[0] * t.r[q].n[1] * t.r[q].n[2] < (int64_t(1) << 31),
                    "%s: a halo region holds 2^31 points or more", who);
            GT355_CHECK(t.blocks < (1 << 30), "%s: too many halo points", who);
            return GT355_OK;
        }

        void add_region(region_table &t, const int32_t lo[3], const int32_t n[3], const int32_t dir[3]) {
            region r{};
            int64_t vol = 1;
            for (int d = 0; d < 3; ++d) {
                r.lo[d] = lo[d], r.n[d] = n[d], r.dir[d] = dir[d];
                vol *= n[d];
            }
            if (vol == 0)
                return;
            r.start = r.src = t.total;
            r.block0 = t.blocks;
            r.d0 = make_fastdiv((uint32_t)r.n[0]);
            r.d1 = make_fastdiv((uint32_t)r.n[0] * (uint32_t)r.n[1]);
            t.blocks += (int)((vol + BC_CHUNK - 1) / BC_CHUNK);
            t.total += vol;
            t.r[t.count++] = r;
        }

        // The halo fill with the i faces as row-gap runs, if the call allows it: the i halo filled, no k halo
        // regions, the logical i box exactly halo + interior + halo, strides of gt355_storage_layout (row padding
        // owned by the layout), 32-byte alignment of every run (first interior column of each field, ni, the
        // strides), at least two rows.  Then the table holds the j faces and the two end rows' i halos of each
        // level, and g the runs between consecutive rows.
        bool plan_gap_fill(const bc_args &a, int64_t es, const int32_t halo[3][5], const int32_t mode[3],
            region_table &t, gap_args &g) {
            const int32_t *hi = halo[0], *hj = halo[1], *hk = halo[2];
            if (mode[0] == 0 || hi[0] == 0 || hi[1] == 0 || (mode[2] != 0 && (hk[0] || hk[1])))
                return false;
            if (hi[2] != hi[0] || hi[4] != hi[3] + hi[1] + 1)
                return false;
            const int64_t align = 128 / es, row = (hi[4] + align - 1) / align * align;
            if (a.s[1] != row || a.s[2] != row * hj[4])
                return false;
            const int64_t ni = hi[3] - hi[2] + 1;
            if ((ni * es) % 32 || (a.s[1] * es) % 32 || (a.s[2] * es) % 32)
                return false;
            for (int f = 0; f < a.nfields; ++f)
                if (((uintptr_t)a.dst[f] + (uintptr_t)(hi[2] * es)) % 32)
                    return false;
            const int row0 = mode[1] ? hj[2] - hj[0] : hj[2], row1 = mode[1] ? hj[3] + hj[1] : hj[3];
            const int nrows = row1 - row0 + 1, nk = hk[3] - hk[2] + 1;
            if (nrows < 2)
                return false;
            const int64_t len = a.s[1] - ni, cpr = len * es / 16;
            const int64_t total = cpr * (nrows - 1) * nk;
            if (total >= (int64_t(1) << 31))
                return false;
            t = region_table{};
            // the j faces (all interior columns of the j halo rows)
            if (mode[1]) {
                for (int dj = -1; dj <= 1; dj += 2) {
                    int32_t lo[3], n[3];
                    const int32_t dir[3] = {0, dj, 0};
                    dim_range(hi, 0, 0, lo[0], n[0]);
                    dim_range(hj, dj, 0, lo[1], n[1]);
                    dim_range(hk, 0, 0, lo[2], n[2]);
                    add_region(t, lo, n, dir);
                }
            }
            // the left halo of the first row and the right halo of the last row of each level
            {
                const int32_t lo[3] = {hi[2] - hi[0], row0, hk[2]}, n[3] = {hi[0], 1, nk};
                const int32_t dir[3] = {-1, row0 < hj[2] ? -1 : 0, 0};
                add_region(t, lo, n, dir);
            }
            {
                const int32_t lo[3] = {hi[3] + 1, row1, hk[2]}, n[3] = {hi[1], 1, nk};
                const int32_t dir[3] = {1, row1 > hj[3] ? 1 : 0, 0};
                add_region(t, lo, n, dir);
            }
            g = gap_args{};
            g.row0 = row0, g.nrows = nrows, g.k0 = hk[2], g.h0 = hi[0], g.h1 = hi[1];
            g.len = (int32_t)len, g.cpr = (int32_t)cpr, g.total = (uint32_t)total;
            g.dcpr = make_fastdiv((uint32_t)cpr);
            g.druns = make_fastdiv((uint32_t)(nrows - 1));
            return true;
        }

        int fill_args(bc_args &a, int dtype, void *const *dst, void *const *src, int nfields, const int32_t halo[3][5],
            const int64_t stride[3], const char *who) {
            GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "%s: dtype %d", who, dtype);
            GT355_CHECK(nfields >= 1 && nfields <= MAX_FIELDS, "%s: nfields %d (1..%d)", who, nfields, MAX_FIELDS);
            GT355_CHECK(dst && stride, "%s: null argument", who);
            GT355_CHECK(stride[0] == 1 && stride[1] >= halo[0][4] && stride[2] >= stride[1] * halo[1][4],
                "%s: strides {%lld,%lld,%lld} do not fit the halo lengths", who, (long long)stride[0],
                (long long)stride[1], (long long)stride[2]);
            a = bc_args{};
            a.nfields = nfields;
            for (int f = 0; f < nfields; ++f) {
                GT355_CHECK(dst[f] != nullptr, "%s: null field %d", who, f);
                a.dst[f] = dst[f];
                a.src[f] = src ? src[f] : nullptr;
            }
            for (int d = 0; d < 3; ++d) {
                a.s[d] = stride[d];
                a.begin[d] = halo[d][2];
                a.end[d] = halo[d][3];
            }
            return GT355_OK;
        }

    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_bc_apply(int kind, int dtype, void *const *dst, void *const *src, int nfields, double value,
    const int32_t halo[3][5], const int64_t stride[3], uint32_t dir_mask, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(halo != nullptr, "gt355_bc_apply: null halo");
    int rc;
    if ((rc = check_halo(halo, "gt355_bc_apply")))
        return rc;
    GT355_CHECK(kind >= 0 && kind <= 2, "gt355_bc_apply: kind %d (0 zero, 1 value, 2 copy)", kind);
    bc_args a;
    if ((rc = fill_args(a, dtype, dst, src, nfields, halo, stride, "gt355_bc_apply")))
        return rc;
    if (kind == 2) {
        GT355_CHECK(src != nullptr, "gt355_bc_apply: copy needs source fields");
        for (int f = 0; f < nfields; ++f)
            GT355_CHECK(src[f] != nullptr, "gt355_bc_apply: null source field %d", f);
    }
    a.value = value;
    region_table t = make_table(halo, dir_mask, 0, nullptr);
    if (t.total == 0)
        return GT355_OK;
    if ((rc = check_table(t, "gt355_bc_apply")))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const int nb = t.blocks;
#define GT355_BC_LAUNCH(T)                                                  \
    switch (kind) {                                                         \
    case 0: bc_kernel<T, 0><<<nb, BC_BLOCK, 0, s>>>(t, a, gap_args{}); break;    \
    case 1: bc_kernel<T, 1><<<nb, BC_BLOCK, 0, s>>>(t, a, gap_args{}); break;    \
    default: bc_kernel<T, 2><<<nb, BC_BLOCK, 0, s>>>(t, a, gap_args{}); break;   \
    }
    if (dtype == GT355_F64) {
        GT355_BC_LAUNCH(double)
    } else {
        GT355_BC_LAUNCH(float)
    }
#undef GT355_BC_LAUNCH
    return check_launch("gt355_bc_apply");
}

extern "C" int gt355_halo_fill(int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], const int32_t mode[3], void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(halo != nullptr && mode != nullptr, "gt355_halo_fill: null argument");
    int rc;
    if ((rc = check_halo(halo, "gt355_halo_fill")))
        return rc;
    bc_args a;
    if ((rc = fill_args(a, dtype, fields, nullptr, nfields, halo, stride, "gt355_halo_fill")))
        return rc;
    for (int d = 0; d < 3; ++d) {
        GT355_CHECK(mode[d] >= 0 && mode[d] <= 2, "gt355_halo_fill: mode[%d] = %d", d, mode[d]);
        int len = halo[d][3] - halo[d][2] + 1;
        GT355_CHECK(mode[d] != 2 || (halo[d][0] <= len && halo[d][1] <= len),
            "gt355_halo_fill: periodic halo wider than the interior in dim %d", d);
        a.mode[d] = mode[d];
    }
    region_table t{};
    gap_args g{};
    // the i faces as whole-sector row-gap runs where the layout allows (GT355_TUNE_halo_gap=0: per point)
    if (!(tune_param("halo_gap", 1) && plan_gap_fill(a, dtype == GT355_F64 ? 8 : 4, halo, mode, t, g))) {
        t = make_table(halo, 0x7FFFFFFu, 0, mode);
        g = gap_args{};
    }
    if (t.total == 0 && g.total == 0)
        return GT355_OK;
    if ((rc = check_table(t, "gt355_halo_fill")))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const int nb = t.blocks + (int)((g.total + BC_BLOCK - 1) / BC_BLOCK);
    if (dtype == GT355_F64)
        bc_kernel<double, 3><<<nb, BC_BLOCK, 0, s>>>(t, a, g);
    else
        bc_kernel<float, 3><<<nb, BC_BLOCK, 0, s>>>(t, a, g);
    return check_launch("gt355_halo_fill");
}

extern "C" int gt355_halo_buffer_size(int dtype, int nfields, const int32_t halo[3][5], uint32_t dir_mask,
    size_t *bytes) {
    GT355_CHECK(halo && bytes, "gt355_halo_buffer_size: null argument");
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_halo_buffer_size: dtype %d", dtype);
    GT355_CHECK(nfields >= 1 && nfields <= MAX_FIELDS, "gt355_halo_buffer_size: nfields %d", nfields);
    int rc;
    if ((rc = check_halo(halo, "gt355_halo_buffer_size")))
        return rc;
    region_table t = make_table(halo, dir_mask, 1, nullptr);
    *bytes = (size_t)t.total * nfields * (dtype == GT355_F64 ? 8 : 4);
    return GT355_OK;
}

static int halo_pack_impl(bool unpack, int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], uint32_t dir_mask, void *buffer, void *stream) {
    const char *who = unpack ? "gt355_halo_unpack" : "gt355_halo_pack";
    GT355_CHECK(halo && buffer, "%s: null argument", who);
    int rc;
    if ((rc = check_halo(halo, who)))
        return rc;
    bc_args a;
    if ((rc = fill_args(a, dtype, fields, nullptr, nfields, halo, stride, who)))
        return rc;
    // pack reads the slab a neighbour in direction d needs; unpack writes our halo on side d, and its segment
    // sizes equal the pack sizes of direction -d, so a buffer received from the neighbour at d fits.
    region_table t = make_table(halo, dir_mask, unpack ? 0 : 1, nullptr);
    if (unpack) {
        // the buffer was sized by gt355_halo_buffer_size (pack volumes); unpack must not read past it
        region_table tp = make_table(halo, dir_mask, 1, nullptr);
        GT355_CHECK(tp.total == t.total,
            "%s: halo volumes (%lld points) differ from the pack volumes (%lld) for this direction mask; use a mask "
            "closed under direction negation or symmetric halos",
            who, (long long)t.total, (long long)tp.total);
    }
    if (t.total == 0)
        return GT355_OK;
    if ((rc = check_table(t, who)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const int nb = t.blocks;
    if (dtype == GT355_F64) {
        if (unpack)
            pack_kernel<double, true><<<nb, BC_BLOCK, 0, s>>>(t, a, (double *)buffer);
        else
            pack_kernel<double, false><<<nb, BC_BLOCK, 0, s>>>(t, a, (double *)buffer);
    } else {
        if (unpack)
            pack_kernel<float, true><<<nb, BC_BLOCK, 0, s>>>(t, a, (float *)buffer);
        else
            pack_kernel<float, false><<<nb, BC_BLOCK, 0, s>>>(t, a, (float *)buffer);
    }
    return check_launch(who);
}

extern "C" int gt355_halo_pack(int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], uint32_t dir_mask, void *buffer, void *stream) {
    GT355_DEFER_FLUSH(stream);
    return halo_pack_impl(false, dtype, fields, nfields, halo, stride, dir_mask, buffer, stream);
}

extern "C" int gt355_halo_unpack(int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], uint32_t dir_mask, const void *buffer, void *stream) {
    GT355_DEFER_FLUSH(stream);
    return halo_pack_impl(true, dtype, fields, nfields, halo, stride, dir_mask, const_cast<void *>(buffer), stream);
}

extern "C" int gt355_halo_exchange_local(int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], uint32_t dir_mask, void *buffer, void *stream) {
    GT355_DEFER_FLUSH(stream);
    const char *who = "gt355_halo_exchange_local";
    GT355_CHECK(halo && buffer, "%s: null argument", who);
    int rc;
    if ((rc = check_halo(halo, who)))
        return rc;
    bc_args a;
    if ((rc = fill_args(a, dtype, fields, nullptr, nfields, halo, stride, who)))
        return rc;
    // the halo of side d receives what the neighbour at d packs for its side -d; on one rank with a periodic
    // process grid that neighbour is this rank, so region d of the unpack reads the segment packed for -d
    const region_table tp = make_table(halo, dir_mask, 1, nullptr);
    region_table tu = make_table(halo, dir_mask, 0, nullptr);
    for (int q = 0; q < tu.count; ++q) {
        region &r = tu.r[q];
        int found = -1;
        for (int p = 0; p < tp.count; ++p)
            if (tp.r[p].dir[0] == -r.dir[0] && tp.r[p].dir[1] == -r.dir[1] && tp.r[p].dir[2] == -r.dir[2])
                found = p;
        GT355_CHECK(found >= 0, "%s: direction (%d,%d,%d) has no opposite in the mask", who, r.dir[0], r.dir[1],
            r.dir[2]);
        const region &o = tp.r[found];
        GT355_CHECK((int64_t)o.n[0] * o.n[1] * o.n[2] == (int64_t)r.n[0] * r.n[1] * r.n[2],
            "%s: halo of direction (%d,%d,%d) and the slab packed for the opposite side differ in size (asymmetric "
            "halo)",
            who, r.dir[0], r.dir[1], r.dir[2]);
        r.src = o.start;
    }
    if (tp.total == 0)
        return GT355_OK;
    if ((rc = check_table(tp, who)) || (rc = check_table(tu, who)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == GT355_F64) {
        pack_kernel<double, false><<<tp.blocks, BC_BLOCK, 0, s>>>(tp, a, (double *)buffer);
        pack_kernel<double, true><<<tu.blocks, BC_BLOCK, 0, s>>>(tu, a, (double *)buffer);
    } else {
        pack_kernel<float, false><<<tp.blocks, BC_BLOCK, 0, s>>>(tp, a, (float *)buffer);
        pack_kernel<float, true><<<tu.blocks, BC_BLOCK, 0, s>>>(tu, a, (float *)buffer);
    }
    return check_launch(who);
}
